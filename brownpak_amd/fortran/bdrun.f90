!> bdrun: Fortran driver of the MI355X BD step.
!>
!> Keeps the reference's entry and call structure: command-line arguments
!> fn_control= and job_tag= (src/main.f90:47-71), the control.txt keys of
!> ctrlpar_t (src/control_m.f90:92-150), the .cfg sections read by
!> read_config (src/config_io_m.f90:247-544), and the bds_run loop
!> (src/bd_solver_m.f90:128-212) with integrate_em / calc_diffusion /
!> calc_drift executed on the GPU through libbdstep (bdstep_c).
!> Only bdintg = EM is available on the device; the Gaussian stream comes
!> from the device Philox generator seeded from job_tag.
program bdrun
use iso_c_binding
use bdstep_c
implicit none

integer, parameter :: rp = c_double
integer, parameter :: mxparam = 12
character(len=512) :: arg, fn_control, job_tag
! control keys (ctrlpar_t)
real(rp) :: rcutoff = 0, tskin = 0, lanc_tol = 0, tim_stp = 0
character(len=8) :: mth_ptgen = 'DIR', mob_fctr = 'CHOL', bdintg = 'EM'
logical :: lhdia = .false., lvdw = .false., lanc_reorth = .false., keep_reuse_noise = .false.
integer :: excluded_atoms = 0, lanc_mxitr = 0, nts_mobsam = 1, precision = 0
integer(c_long_long) :: nts_sim = 0, nts_log = 0, nts_samp = 0
character(len=512) :: fn_cfg = 'config.cfg', fn_out = 'final.xyz'
! configuration (atmcfg_t)
integer :: imcon = 0, natom_types = 1, num_atoms = 0, num_bonds = 0, nbond_types = 0, nvdw = 0
real(rp) :: box(3) = 0
integer, allocatable :: atoms(:), bonds(:,:), bond_styles(:), vdw_styles(:)
real(rp), allocatable :: coordinates(:,:), bond_params(:,:), vdw_params(:,:)
integer :: flow_style = 0
real(rp) :: flow_rate = 0
! run
type(c_ptr) :: ctx
type(bd_opts_t) :: opts
type(bd_runinfo_t) :: info
integer(c_int) :: rc
integer(c_long_long) :: nts, chunk, used
integer :: i, iarg
integer(8) :: t0, t1, rate
real(rp) :: rgsq

fn_control = 'control.txt'; job_tag = ''
do iarg = 1, command_argument_count()
    call get_command_argument(iarg, arg)
    if (index(arg, 'fn_control=') == 1) fn_control = arg(12:)
    if (index(arg, 'job_tag=') == 1) job_tag = arg(9:)
end do

call read_control(trim(fn_control))
call read_cfg(trim(fn_cfg))
select case (trim(bdintg))
case ('EM'); opts%integrator = BD_EM
case ('PC'); opts%integrator = BD_PC
case default
    print '(a)', 'bdrun: bdintg must be EM or PC on the device'
    stop 2
end select

opts%lhdia = merge(1, 0, lhdia)
select case (trim(mob_fctr))
case ('LANC'); opts%mob_fctr = BD_LANC
case ('CHLT'); opts%mob_fctr = BD_CHOL_LT
case default; opts%mob_fctr = BD_CHOL
end select
opts%nts_mobsam = max(nts_mobsam, 1)
opts%lanc_mxitr = lanc_mxitr
opts%lanc_tol = lanc_tol
opts%tim_stp = tim_stp
opts%precision = precision
opts%keep_reuse_noise = merge(1, 0, keep_reuse_noise)
opts%lanc_reorth = merge(1, 0, lanc_reorth)
opts%flow_style = flow_style
opts%flow_rate = flow_rate

ctx = bd_ctx_create(int(num_atoms, c_int), opts)
if (.not. c_associated(ctx)) call fail(-1)
rc = bd_set_box(ctx, int(imcon, c_int), box); if (rc /= BD_OK) call fail(rc)
rc = bd_set_topology(ctx, int(num_bonds, c_int), bonds, int(nbond_types, c_int), bond_styles, bond_params, &
        int(natom_types, c_int), atoms, int(nvdw, c_int), vdw_styles, vdw_params, merge(1, 0, lvdw), &
        int(excluded_atoms, c_int), pt_code(mth_ptgen), rcutoff, tskin)
if (rc /= BD_OK) call fail(rc)
rc = bd_set_coordinates(ctx, coordinates); if (rc /= BD_OK) call fail(rc)
rc = bd_set_rng(ctx, int(seed_from(job_tag), c_int64_t), 0_c_int64_t); if (rc /= BD_OK) call fail(rc)

call system_clock(t0, rate)
nts = 0
print '(a,i0)', 'nts = ', nts
if (nts_log <= 0) nts_log = max(nts_sim, 1_c_long_long)
do while (nts < nts_sim)
    chunk = min(nts_log, nts_sim - nts)
    rc = bd_run(ctx, int(chunk, c_int), c_null_ptr, 0_c_long_long, used)
    if (rc /= BD_OK .and. rc /= BD_WARN_NOTCONV) call fail(rc)
    nts = nts + chunk
    rc = bd_rgsq(ctx, rgsq)
    print '(a,i0,a,es14.6)', 'nts = ', nts, '  rgsq = ', rgsq
end do
call system_clock(t1)
rc = bd_get_runinfo(ctx, info)
print '(a,es12.4)', 'execution time(s) ', real(t1 - t0, rp) / real(rate, rp)
print '(a,i0,a,i0)', 'mobility updates ', info%n_mob_updates, '  Lanczos iterations ', info%total_lanczos_iters
rc = bd_get_coordinates(ctx, coordinates)
open(newunit=i, file=trim(fn_out)//trim(job_tag), status='replace', action='write')
write(i, '(i0)') num_atoms
write(i, '(a,i0)') 'nts = ', nts
do iarg = 1, num_atoms
    write(i, '(a,3es24.16)') 'M ', coordinates(:, iarg)
end do
close(i)
call bd_ctx_destroy(ctx)

contains

subroutine fail(code)
    integer(c_int), intent(in) :: code
    print '(a,i0,2a)', 'bdrun: error ', code, ': ', bd_error_message()
    stop 1
end subroutine

function pt_code(m) result(c)
    character(len=*), intent(in) :: m
    integer(c_int) :: c
    select case (trim(m))
    case ('VER'); c = BD_PT_VER
    case ('CL'); c = BD_PT_CL
    case default; c = BD_PT_DIR
    end select
end function

function seed_from(tag) result(s)
    character(len=*), intent(in) :: tag
    integer(8) :: s
    integer :: k
    s = 88172645463325252_8
    do k = 1, len_trim(tag)
        s = ieor(s * 31_8, int(ichar(tag(k:k)), 8))
    end do
end function

!> Whitespace key-value control file; unknown keys are ignored
!> (control_m.f90:92-150).  Extra device keys: precision, lanc_reorth,
!> keep_reuse_noise, fn_out.
subroutine read_control(fn)
    character(len=*), intent(in) :: fn
    character(len=512) :: line, key, val
    integer :: fu, ios, p
    real(rp) :: d
    open(newunit=fu, file=fn, action='read', status='old')
    do
        read(fu, '(a)', iostat=ios) line
        if (ios /= 0) exit
        p = index(line, '#'); if (p > 0) line = line(:p-1)
        line = adjustl(line)
        if (len_trim(line) == 0) cycle
        p = index(line, ' ')
        key = line(:p-1); val = adjustl(line(p:))
        select case (trim(key))
        case ('rcutoff'); read(val, *) rcutoff
        case ('tskin'); read(val, *) tskin
        case ('mth_ptgen'); read(val, *) mth_ptgen
        case ('lhdia'); read(val, *) lhdia
        case ('lvdw'); read(val, *) lvdw
        case ('excluded_atoms'); read(val, *) excluded_atoms
        case ('mob_fctr'); read(val, *) mob_fctr
        case ('lanc_mxitr'); read(val, *) lanc_mxitr
        case ('lanc_tol'); read(val, *) lanc_tol
        case ('bdintg'); read(val, *) bdintg
        case ('tim_stp'); read(val, *) tim_stp
        case ('nts_sim'); read(val, *) d; nts_sim = int(d, c_long_long)
        case ('nts_mobsam'); read(val, *) d; nts_mobsam = int(d)
        case ('nts_log'); read(val, *) d; nts_log = int(d, c_long_long)
        case ('nts_samp'); read(val, *) d; nts_samp = int(d, c_long_long)
        case ('fn_cfg'); fn_cfg = trim(val)
        case ('fn_out'); fn_out = trim(val)
        case ('precision'); read(val, *) precision
        case ('lanc_reorth'); read(val, *) lanc_reorth
        case ('keep_reuse_noise'); read(val, *) keep_reuse_noise
        end select
    end do
    close(fu)
end subroutine

!> .cfg reader for the sections the BD step uses (config_io_m.f90:247-544).
subroutine read_cfg(fn)
    character(len=*), intent(in) :: fn
    character(len=512) :: line, word
    integer :: fu, ios, k, n, npar, p, t1, t2
    real(rp) :: mat(3,3)
    character(len=32) :: name
    open(newunit=fu, file=fn, action='read', status='old')
    do
        read(fu, '(a)', iostat=ios) line
        if (ios /= 0) exit
        p = index(line, '#'); if (p > 0) line = line(:p-1)
        line = adjustl(line)
        if (len_trim(line) == 0) cycle
        read(line, *) word
        select case (trim(word))
        case ('SIMBOX')
            read(fu, *) mat(:, 1); read(fu, *) mat(:, 2); read(fu, *) mat(:, 3)
            box = [mat(1,1), mat(2,2), mat(3,3)]
        case ('IMCON')
            read(line, *) word, imcon
        case ('ATOM_TYPES')
            read(line, *) word, natom_types
            do k = 1, natom_types
                read(fu, *) name
            end do
        case ('ATOMS')
            read(line, *) word, num_atoms
            allocate(atoms(num_atoms), coordinates(3, num_atoms))
            do k = 1, num_atoms
                read(fu, *) atoms(k), mat(1,1), coordinates(:, k)
            end do
        case ('BOND_TYPES')
            read(line, *) word, nbond_types
            allocate(bond_styles(nbond_types), bond_params(mxparam, nbond_types))
            bond_params = 0
            do k = 1, nbond_types
                read(fu, '(a)') line
                read(line, *) bond_styles(k), npar
                if (npar > 0) read(line, *) t1, t2, bond_params(1:npar, k)
            end do
        case ('BONDS')
            read(line, *) word, num_bonds
            allocate(bonds(3, num_bonds))
            do k = 1, num_bonds
                read(fu, *) bonds(:, k)
            end do
        case ('VDW')
            read(line, *) word, nvdw
            allocate(vdw_styles(nvdw), vdw_params(mxparam, nvdw))
            vdw_params = 0
            do k = 1, nvdw
                read(fu, '(a)') line
                read(line, *) t1, t2, vdw_styles(k), npar
                if (npar > 0) read(line, *) t1, t2, n, n, vdw_params(1:npar, k)
            end do
        case ('FLOW_FIELD')
            read(fu, '(a)') line
            read(line, *) flow_style
            if (flow_style > 0) read(line, *) flow_style, npar, flow_rate
        end select
    end do
    close(fu)
    if (.not. allocated(bonds)) allocate(bonds(3, 1), bond_styles(1), bond_params(mxparam, 1))
    if (.not. allocated(vdw_styles)) allocate(vdw_styles(1), vdw_params(mxparam, 1))
    if (num_bonds == 0) then
        bonds = 0; bond_styles = 0; bond_params = 0
    end if
end subroutine

end program bdrun
