!> bdrun_io: the reference's input and output files for the bdrun driver.
!>
!>   read_config    .cfg reader (src/config_io_m.f90:247-544), readline
!>                  semantics of src/strings_m.f90:928-949 (blank and
!>                  comment lines skipped, trailing comments stripped)
!>   write_dump     revive file (src/config_io_m.f90:176-243); after the
!>                  reference's last item it appends the device generator's
!>                  state: 8 bytes 'BDPHILOX', int64 seed, int64 offset.
!>                  The reference's read_dump stops before that record.
!>   read_dump      (src/config_io_m.f90:15-172) plus that record if present
!>   traj_*         trajectory file with coordinate frames
!>                  (src/trajectory_m.f90:84-135, 140-175, 283-331)
!>   stats_*        statistics file header and records
!>                  (src/stats_m.f90:53-105, 167-294)
!> Unformatted files are stream access: every item is written as its raw
!> bytes in order (default integer 4 bytes, real(rp) 8, integer(ip_long) 8,
!> character(len=8) 8).
module bdrun_io
use iso_c_binding
implicit none
private

integer, parameter :: rp = c_double
integer, parameter, public :: mxparam = 12
character(len=8), parameter :: rng_tag = 'BDPHILOX'

type, public :: cfg_t
    integer :: imcon = 0
    real(rp) :: basis(3,3) = 0.0_rp
    integer :: num_atom_types = 0
    character(len=8), allocatable :: atom_names(:)
    integer, allocatable :: atom_styles(:)
    real(rp), allocatable :: atom_mass(:)
    integer :: num_atoms = 0
    integer, allocatable :: atoms(:)
    real(rp), allocatable :: charge(:)
    real(rp), allocatable :: coordinates(:,:)
    integer :: num_bond_types = 0
    integer, allocatable :: bond_styles(:)
    real(rp), allocatable :: bond_params(:,:)
    integer :: num_bonds = 0
    integer, allocatable :: bonds(:,:)
    integer :: num_angle_types = 0
    integer, allocatable :: angle_styles(:)
    real(rp), allocatable :: angle_params(:,:)
    integer :: num_angles = 0
    integer, allocatable :: angles(:,:)
    integer :: num_dihedral_types = 0
    integer, allocatable :: dihedral_styles(:)
    real(rp), allocatable :: dihedral_params(:,:)
    integer :: num_dihedrals = 0
    integer, allocatable :: dihedrals(:,:)
    integer :: num_branches = 0
    integer, allocatable :: branches(:,:)
    integer :: num_molecule_types = 0
    character(len=8), allocatable :: molecule_names(:)
    integer, allocatable :: molecule_pop(:)
    integer :: num_molecules = 0
    integer, allocatable :: molecules(:,:)
    real(rp) :: molecule_com(3) = 0.0_rp
    integer :: num_tether_types = 0
    integer, allocatable :: tether_styles(:)
    real(rp), allocatable :: tether_params(:,:)
    integer :: num_tethers = 0
    integer, allocatable :: tethers(:,:)
    real(rp), allocatable :: tether_points(:,:)
    integer :: num_vdw_types = 0
    integer, allocatable :: vdw_styles(:)
    integer, allocatable :: vdw_pairs(:,:)
    real(rp), allocatable :: vdw_params(:,:)
    integer :: num_externals = 0
    integer, allocatable :: external_styles(:)
    real(rp), allocatable :: external_params(:,:)
    integer :: flow_style = 0
    real(rp), allocatable :: flow_params(:)
end type

!> An open trajectory file (trajectory_t, src/trajectory_m.f90:33-45).
type, public :: traj_t
    integer :: header_size = 0
    integer :: frame_size = 0
    integer :: num_atoms = 0
    integer :: frmcmp(4) = 0
    integer :: unit = 0
    integer :: num_frames = 0
    logical :: isopen = .false.
end type

public :: read_config, write_dump, read_dump
public :: traj_create, traj_open, traj_append, traj_close
public :: stats_open, stats_write_line

contains

!------------------------------------------------------------------ .cfg

subroutine readline(fu, line, ios)
    integer, intent(in) :: fu
    character(len=*), intent(out) :: line
    integer, intent(out) :: ios
    integer :: p
    character(len=len(line)) :: s
    do
        read(fu, '(a)', iostat=ios) line
        if (ios /= 0) return
        s = adjustl(line)
        if (len_trim(s) == 0) cycle
        if (s(1:1) == '#') cycle
        p = index(line, '#')
        if (p > 0) line = line(:p-1)
        return
    end do
end subroutine

!> The text after the first blank-separated word of line.
function after_word(line) result(rest)
    character(len=*), intent(in) :: line
    character(len=len(line)) :: rest
    character(len=len(line)) :: s
    integer :: p
    s = adjustl(line)
    p = index(s, ' ')
    if (p == 0) then
        rest = ''
    else
        rest = adjustl(s(p:))
    end if
end function

logical function starts(line, key)
    character(len=*), intent(in) :: line, key
    starts = .false.
    if (len(line) >= len(key)) starts = (line(1:len(key)) == key)
end function

!> A style/parameter row "style npar p(1:npar)" after nlead leading integers
!> (read into lead).
subroutine read_param_row(fu, nlead, lead, style, params)
    integer, intent(in) :: fu, nlead
    integer, intent(out) :: lead(:), style
    real(rp), intent(out) :: params(:)
    character(len=1024) :: line
    integer :: ios, npar, k, sty2
    integer :: dum(4)
    call readline(fu, line, ios)
    read(line, *) (lead(k), k = 1, nlead), style, npar
    params = 0.0_rp
    if (npar > 0) read(line, *) (dum(k), k = 1, nlead), sty2, npar, params(1:npar)
end subroutine

subroutine read_config(fn, c)
    character(len=*), intent(in) :: fn
    type(cfg_t), intent(out) :: c
    character(len=1024) :: line, rest
    real(rp) :: mat(3,3)
    integer :: fu, ios, i, n, nt, lead(2), npar, sty
    open(newunit=fu, file=fn, action='read', status='old')
    do
        call readline(fu, line, ios)
        if (ios /= 0) exit
        line = adjustl(line)
        rest = after_word(line)
        if (starts(line, 'SIMBOX')) then
            if (len_trim(rest) == 0) then
                read(fu, *) mat(:,1)
                read(fu, *) mat(:,2)
                read(fu, *) mat(:,3)
                read(fu, *)
                call readline(fu, line, ios)
                rest = after_word(line)
                read(rest, *) c%imcon
            else
                read(rest, *) c%imcon
                read(fu, *) mat(:,1)
                read(fu, *) mat(:,2)
                read(fu, *) mat(:,3)
            end if
            c%basis = mat
        else if (starts(line, 'ATOM_TYPES')) then
            read(rest, *) nt
            c%num_atom_types = nt
            allocate(c%atom_names(nt), c%atom_styles(nt), c%atom_mass(nt))
            do i = 1, nt
                read(fu, *) c%atom_names(i), c%atom_styles(i), c%atom_mass(i)
            end do
        else if (starts(line, 'ATOMS')) then
            read(rest, *) n
            c%num_atoms = n
            allocate(c%atoms(n), c%charge(n), c%coordinates(3,n))
            do i = 1, n
                read(fu, *) c%atoms(i), c%charge(i), c%coordinates(:,i)
            end do
        else if (starts(line, 'BOND_TYPES')) then
            read(rest, *) nt
            c%num_bond_types = nt
            allocate(c%bond_styles(nt), c%bond_params(mxparam,nt))
            do i = 1, nt
                call read_param_row(fu, 0, lead, c%bond_styles(i), c%bond_params(:,i))
            end do
        else if (starts(line, 'BONDS')) then
            read(rest, *) n
            c%num_bonds = n
            allocate(c%bonds(3,n))
            do i = 1, n
                read(fu, *) c%bonds(:,i)
            end do
        else if (starts(line, 'ANGLE_TYPES')) then
            read(rest, *) nt
            c%num_angle_types = nt
            allocate(c%angle_styles(nt), c%angle_params(mxparam,nt))
            do i = 1, nt
                call read_param_row(fu, 0, lead, c%angle_styles(i), c%angle_params(:,i))
            end do
        else if (starts(line, 'ANGLES')) then
            read(rest, *) n
            c%num_angles = n
            allocate(c%angles(4,n))
            do i = 1, n
                read(fu, *) c%angles(:,i)
            end do
        else if (starts(line, 'DIHEDRAL_TYPES')) then
            read(rest, *) nt
            c%num_dihedral_types = nt
            allocate(c%dihedral_styles(nt), c%dihedral_params(mxparam,nt))
            do i = 1, nt
                call read_param_row(fu, 0, lead, c%dihedral_styles(i), c%dihedral_params(:,i))
            end do
        else if (starts(line, 'DIHEDRALS')) then
            read(rest, *) n
            c%num_dihedrals = n
            allocate(c%dihedrals(5,n))
            do i = 1, n
                read(fu, *) c%dihedrals(:,i)
            end do
        else if (starts(line, 'BRANCHES')) then
            read(rest, *) n
            c%num_branches = n
            allocate(c%branches(3,n))
            do i = 1, n
                read(fu, *) c%branches(:,i)
            end do
        else if (starts(line, 'MOLECULE_TYPES')) then
            read(rest, *) nt
            c%num_molecule_types = nt
            allocate(c%molecule_names(nt), c%molecule_pop(nt))
            do i = 1, nt
                read(fu, *) c%molecule_names(i), c%molecule_pop(i)
            end do
        else if (starts(line, 'MOLECULES')) then
            read(rest, *) n
            c%num_molecules = n
            allocate(c%molecules(9,n))
            do i = 1, n
                read(fu, *) c%molecules(:,i)
            end do
        else if (starts(line, 'TETHER_TYPES')) then
            read(rest, *) nt
            c%num_tether_types = nt
            allocate(c%tether_styles(nt), c%tether_params(mxparam,nt))
            do i = 1, nt
                call read_param_row(fu, 0, lead, c%tether_styles(i), c%tether_params(:,i))
            end do
        else if (starts(line, 'TETHERS')) then
            read(rest, *) n
            c%num_tethers = n
            allocate(c%tethers(2,n), c%tether_points(3,n))
            do i = 1, n
                read(fu, *) c%tethers(:,i), c%tether_points(:,i)
            end do
        else if (starts(line, 'VDW')) then
            read(rest, *) nt
            c%num_vdw_types = nt
            allocate(c%vdw_styles(nt), c%vdw_pairs(2,nt), c%vdw_params(mxparam,nt))
            do i = 1, nt
                call read_param_row(fu, 2, lead, c%vdw_styles(i), c%vdw_params(:,i))
                c%vdw_pairs(:,i) = lead
            end do
        else if (starts(line, 'EXTERNAL')) then
            read(rest, *) n
            c%num_externals = n
            allocate(c%external_styles(n), c%external_params(mxparam,n))
            do i = 1, n
                call read_param_row(fu, 0, lead, c%external_styles(i), c%external_params(:,i))
            end do
        else if (starts(line, 'FLOW_FIELD')) then
            allocate(c%flow_params(mxparam))
            c%flow_params = 0.0_rp
            call readline(fu, line, ios)
            read(line, *) c%flow_style
            if (c%flow_style > 0) then
                read(line, *) sty, npar
                if (npar > 0) read(line, *) sty, npar, c%flow_params(1:npar)
            end if
        end if
    end do
    close(fu)
    call fill_defaults(c)
end subroutine

!> Arrays the dump writes unconditionally exist even when the .cfg has no
!> such section (the reference would write unallocated arrays as empty).
subroutine fill_defaults(c)
    type(cfg_t), intent(in out) :: c
    if (.not. allocated(c%atom_names)) allocate(c%atom_names(0), c%atom_styles(0), c%atom_mass(0))
    if (.not. allocated(c%atoms)) allocate(c%atoms(0), c%charge(0), c%coordinates(3,0))
    if (.not. allocated(c%molecule_names)) allocate(c%molecule_names(0), c%molecule_pop(0))
    if (.not. allocated(c%molecules)) allocate(c%molecules(9,0))
end subroutine

!------------------------------------------------------------------ dump

subroutine write_dump(fn, nts, c, seed, offset)
    character(len=*), intent(in) :: fn
    integer(c_int64_t), intent(in) :: nts, seed, offset
    type(cfg_t), intent(in) :: c
    integer :: fu
    open(newunit=fu, file=fn, access='stream', form='unformatted', action='write', status='replace')
    write(fu) nts
    write(fu) c%imcon
    write(fu) c%basis
    write(fu) c%num_atom_types
    write(fu) c%atom_names, c%atom_styles, c%atom_mass
    write(fu) c%num_atoms
    write(fu) c%atoms, c%coordinates, c%charge
    write(fu) c%num_bond_types
    if (c%num_bond_types > 0) write(fu) c%bond_styles, c%bond_params
    write(fu) c%num_bonds
    if (c%num_bonds > 0) write(fu) c%bonds
    write(fu) c%num_angle_types
    if (c%num_angle_types > 0) write(fu) c%angle_styles, c%angle_params
    write(fu) c%num_angles
    if (c%num_angles > 0) write(fu) c%angles
    write(fu) c%num_dihedral_types
    if (c%num_dihedral_types > 0) write(fu) c%dihedral_styles, c%dihedral_params
    write(fu) c%num_dihedrals
    if (c%num_dihedrals > 0) write(fu) c%dihedrals
    write(fu) c%num_branches
    if (c%num_branches > 0) write(fu) c%branches
    write(fu) c%num_molecule_types
    write(fu) c%molecule_names, c%molecule_pop
    write(fu) c%num_molecules
    write(fu) c%molecules
    write(fu) c%num_tether_types
    if (c%num_tether_types > 0) write(fu) c%tether_styles, c%tether_params
    write(fu) c%num_tethers
    if (c%num_tethers > 0) write(fu) c%tethers, c%tether_points
    write(fu) c%num_vdw_types
    if (c%num_vdw_types > 0) write(fu) c%vdw_pairs, c%vdw_styles, c%vdw_params
    write(fu) c%num_externals
    if (c%num_externals > 0) write(fu) c%external_styles, c%external_params
    write(fu) c%flow_style
    if (c%flow_style /= 0) write(fu) c%flow_params
    if ((c%num_tethers == 0) .and. (c%imcon == 0)) write(fu) c%molecule_com
    write(fu) rng_tag, seed, offset
    close(fu)
end subroutine

!> has_rng: the file ends with the generator record (seed, offset set).
subroutine read_dump(fn, nts, c, has_rng, seed, offset)
    character(len=*), intent(in) :: fn
    integer(c_int64_t), intent(out) :: nts, seed, offset
    type(cfg_t), intent(out) :: c
    logical, intent(out) :: has_rng
    character(len=8) :: tag
    integer :: fu, n, nt, ios
    open(newunit=fu, file=fn, access='stream', form='unformatted', action='read', status='old')
    read(fu) nts
    read(fu) c%imcon
    read(fu) c%basis
    read(fu) nt
    c%num_atom_types = nt
    allocate(c%atom_names(nt), c%atom_styles(nt), c%atom_mass(nt))
    read(fu) c%atom_names, c%atom_styles, c%atom_mass
    read(fu) n
    c%num_atoms = n
    allocate(c%atoms(n), c%coordinates(3,n), c%charge(n))
    read(fu) c%atoms, c%coordinates, c%charge
    read(fu) nt
    c%num_bond_types = nt
    if (nt > 0) then
        allocate(c%bond_styles(nt), c%bond_params(mxparam,nt))
        read(fu) c%bond_styles, c%bond_params
    end if
    read(fu) n
    c%num_bonds = n
    if (n > 0) then
        allocate(c%bonds(3,n))
        read(fu) c%bonds
    end if
    read(fu) nt
    c%num_angle_types = nt
    if (nt > 0) then
        allocate(c%angle_styles(nt), c%angle_params(mxparam,nt))
        read(fu) c%angle_styles, c%angle_params
    end if
    read(fu) n
    c%num_angles = n
    if (n > 0) then
        allocate(c%angles(4,n))
        read(fu) c%angles
    end if
    read(fu) nt
    c%num_dihedral_types = nt
    if (nt > 0) then
        allocate(c%dihedral_styles(nt), c%dihedral_params(mxparam,nt))
        read(fu) c%dihedral_styles, c%dihedral_params
    end if
    read(fu) n
    c%num_dihedrals = n
    if (n > 0) then
        allocate(c%dihedrals(5,n))
        read(fu) c%dihedrals
    end if
    read(fu) n
    c%num_branches = n
    if (n > 0) then
        allocate(c%branches(3,n))
        read(fu) c%branches
    end if
    read(fu) nt
    c%num_molecule_types = nt
    allocate(c%molecule_names(nt), c%molecule_pop(nt))
    read(fu) c%molecule_names, c%molecule_pop
    read(fu) n
    c%num_molecules = n
    allocate(c%molecules(9,n))
    read(fu) c%molecules
    read(fu) nt
    c%num_tether_types = nt
    if (nt > 0) then
        allocate(c%tether_styles(nt), c%tether_params(mxparam,nt))
        read(fu) c%tether_styles, c%tether_params
    end if
    read(fu) n
    c%num_tethers = n
    if (n > 0) then
        allocate(c%tethers(2,n), c%tether_points(3,n))
        read(fu) c%tethers, c%tether_points
    end if
    read(fu) nt
    c%num_vdw_types = nt
    if (nt > 0) then
        allocate(c%vdw_styles(nt), c%vdw_pairs(2,nt), c%vdw_params(mxparam,nt))
        read(fu) c%vdw_pairs, c%vdw_styles, c%vdw_params
    end if
    read(fu) n
    c%num_externals = n
    if (n > 0) then
        allocate(c%external_styles(n), c%external_params(mxparam,n))
        read(fu) c%external_styles, c%external_params
    end if
    read(fu) c%flow_style
    if (c%flow_style /= 0) then
        allocate(c%flow_params(mxparam))
        read(fu) c%flow_params
    end if
    if ((c%num_tethers == 0) .and. (c%imcon == 0)) read(fu) c%molecule_com
    has_rng = .false.
    seed = 0
    offset = 0
    read(fu, iostat=ios) tag
    if (ios == 0 .and. tag == rng_tag) then
        read(fu, iostat=ios) seed, offset
        has_rng = (ios == 0)
    end if
    close(fu)
end subroutine

!------------------------------------------------------------ trajectory

!> traj_create with frmcmp = (1, 0, 0, 0): coordinates only.
subroutine traj_create(fn, na, t)
    character(len=*), intent(in) :: fn
    integer, intent(in) :: na
    type(traj_t), intent(out) :: t
    t%num_atoms = na
    t%frmcmp = [1, 0, 0, 0]
    t%frame_size = 8 + 3 * na * 8
    t%header_size = 7 * 4
    open(newunit=t%unit, file=fn, access='stream', form='unformatted', action='readwrite', status='replace')
    t%isopen = .true.
    t%num_frames = 0
    write(t%unit) t%header_size
    write(t%unit) t%frame_size
    write(t%unit) t%num_atoms, t%frmcmp
end subroutine

subroutine traj_open(fn, t)
    character(len=*), intent(in) :: fn
    type(traj_t), intent(out) :: t
    integer(c_int64_t) :: file_size
    open(newunit=t%unit, file=fn, access='stream', form='unformatted', action='readwrite', status='old')
    t%isopen = .true.
    inquire(unit=t%unit, size=file_size)
    read(t%unit) t%header_size
    read(t%unit) t%frame_size
    read(t%unit) t%num_atoms, t%frmcmp
    t%num_frames = int((file_size - t%header_size) / t%frame_size)
end subroutine

subroutine traj_append(t, nts, crd)
    type(traj_t), intent(in out) :: t
    integer(c_int64_t), intent(in) :: nts
    real(rp), intent(in) :: crd(:,:)
    integer(c_int64_t) :: offset
    offset = int(t%num_frames, c_int64_t) * t%frame_size + t%header_size + 1
    write(t%unit, pos=offset) nts
    if (t%num_atoms /= 0) write(t%unit) crd(:, 1:t%num_atoms)
    t%num_frames = t%num_frames + 1
end subroutine

subroutine traj_close(t)
    type(traj_t), intent(in out) :: t
    if (t%isopen) close(t%unit)
    t%isopen = .false.
end subroutine

!------------------------------------------------------------------ stats

!> stats_init's file handling: a new run replaces the file and writes the
!> header; a revived run appends to an existing file.
subroutine stats_open(fn, revive, c, lvdw, fu)
    character(len=*), intent(in) :: fn
    logical, intent(in) :: revive, lvdw
    type(cfg_t), intent(in) :: c
    integer, intent(out) :: fu
    logical :: lexists
    if (revive) then
        inquire(file=fn, exist=lexists)
        if (lexists) then
            open(newunit=fu, file=fn, action='write', position='append', status='old')
            return
        end if
        open(newunit=fu, file=fn, action='write', status='new')
    else
        open(newunit=fu, file=fn, action='write', status='replace')
    end if
    write(fu, '(a12)', advance='no') 'nts'
    if (c%num_bonds > 0) write(fu, 200, advance='no') 'bndlen_avg', 'bndlen_min', 'bndlen_max', 'enrg_bnd'
    if (c%num_angles > 0) write(fu, 100, advance='no') 'enrg_ang'
    if (c%num_dihedrals > 0) write(fu, 100, advance='no') 'enrg_dhd'
    if (lvdw) write(fu, 100, advance='no') 'enrg_vdw'
    if (c%num_tethers > 0) write(fu, 100, advance='no') 'enrg_teth'
    if (c%num_externals > 0) write(fu, 100, advance='no') 'enrg_ext'
    write(fu, 100, advance='no') 'enrg_tot'
    if (c%num_atoms > 0) write(fu, 200, advance='no') 'sxx', 'syx', 'szx', 'syy', 'szy', 'szz'
    if (c%num_bonds > 0) write(fu, 200, advance='no') 'rgsq', 'reedsq', 'spanx', 'spany', 'spanz'
    if ((c%imcon == 0) .and. (c%num_bonds > 0)) then
        write(fu, 200, advance='no') 'reevx', 'reevy', 'reevz'
        if (c%flow_style == 0) write(fu, 200, advance='no') 'comx', 'comy', 'comz'
    end if
    write(fu, *)
100 format(1x,a14)
200 format(*(1x,a14))
end subroutine

!> One stats_write record from bd_stats_get's output s(30): bndlen(3),
!> energies(7: bond, angle, dihedral, vdw, tether, external, total),
!> stress(9), rgsq, reedsq, span(3), reev(3), molecule_com(3).
subroutine stats_write_line(fu, nts, c, lvdw, s)
    integer, intent(in) :: fu
    integer(c_int64_t), intent(in) :: nts
    type(cfg_t), intent(in) :: c
    logical, intent(in) :: lvdw
    real(rp), intent(in) :: s(30)
    write(fu, '(i12)', advance='no') nts
    if (c%num_bonds > 0) write(fu, 200, advance='no') s(1), s(2), s(3), s(4)
    if (c%num_angles > 0) write(fu, 100, advance='no') s(5)
    if (c%num_dihedrals > 0) write(fu, 100, advance='no') s(6)
    if (lvdw) write(fu, 100, advance='no') s(7)
    if (c%num_tethers > 0) write(fu, 100, advance='no') s(8)
    if (c%num_externals > 0) write(fu, 100, advance='no') s(9)
    write(fu, 100, advance='no') s(10)
    ! stress_accu(1,1), (2,1), (3,1), (2,2), (3,2), (3,3)
    if (c%num_atoms > 0) write(fu, 200, advance='no') s(11), s(12), s(13), s(15), s(16), s(19)
    if (c%num_bonds > 0) write(fu, 200, advance='no') s(20), s(21), s(22:24)
    if ((c%imcon == 0) .and. (c%num_bonds > 0)) then
        write(fu, 200, advance='no') s(25:27)
        if (c%flow_style == 0) write(fu, 200, advance='no') s(28:30)
    end if
    write(fu, *)
    flush(fu)
100 format(1x,es14.6)
200 format(*(1x,es14.6))
end subroutine

end module bdrun_io
