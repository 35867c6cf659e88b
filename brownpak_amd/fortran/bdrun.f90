!> bdrun: Fortran driver of the MI355X BD step.
!>
!> Keeps the reference's entry and call structure: command-line arguments
!> fn_control= and job_tag= (src/main.f90:47-71), the control.txt keys of
!> ctrlpar_t (src/control_m.f90:92-150), run set-up (src/setup_m.f90:39-125:
!> read_config or, with lrevive, read_dump; stats_init; the trajectory;
!> the random seed) and the bds_run loop (src/bd_solver_m.f90:128-212):
!> centre an isolated molecule, dump the starting configuration, then steps
!> until nts > nts_sim, with statistics accumulated every step, a log line
!> every nts_log steps, a revive dump every nts_dump, a statistics record
!> and a trajectory frame every nts_samp, and a final dump.  integrate_em /
!> calc_diffusion / calc_drift run on the GPU through libbdstep (bdstep_c);
!> the driver calls bd_run for the steps between two of those events.
!> bdintg EM, SE (semi-implicit, NITSOL keys se_nlmxitr, se_kdmax, se_tol)
!> and PC (predictor-corrector) are available on the device.
!>
!> The Gaussian stream is the device Philox generator.  Its seed comes from
!> random_seed.txt<job_tag> with read_seed T (init_stream,
!> src/random_m.f90:57-86), else from job_tag; write_seed T writes it there.
!> Every revive dump ends with the generator's (seed, offset), and a revived
!> run continues that stream without centring the molecule again, so a run
!> split by a revive at a multiple of nts_mobsam takes the same steps as an
!> uninterrupted one.  An interval key of 0 (nts_log,
!> nts_dump, nts_samp) disables that event.
program bdrun
use iso_c_binding
use bdstep_c
use bdrun_io
implicit none

integer, parameter :: rp = c_double
character(len=512) :: arg, fn_control, job_tag
! control keys (ctrlpar_t)
real(rp) :: rcutoff = 0, tskin = 0, lanc_tol = 0, tim_stp = 0
character(len=8) :: mth_ptgen = 'DIR', mob_fctr = 'CHOL', bdintg = 'EM'
logical :: lhdia = .false., lvdw = .false., lanc_reorth = .false., keep_reuse_noise = .false.
logical :: lrevive = .false., read_seed = .false., write_seed = .false., write_traj = .false.
integer :: excluded_atoms = 0, lanc_mxitr = 0, nts_mobsam = 1, precision = 0
integer :: stats_binsize(2) = 0, se_nlmxitr = 0, se_kdmax = 0
real(rp) :: se_tol(2) = 0
integer(c_int64_t) :: nts_sim = 0, nts_log = 0, nts_samp = 0, nts_dump = 0
character(len=512) :: fn_cfg = 'config.cfg', fn_out = 'final.xyz', fn_revive = 'revive.bin'
character(len=512) :: fn_stats = 'stats.txt', fn_traj = 'traj.bin', fn_xi = ''
character(len=8) :: device_path = 'run'
! configuration (atmcfg_t) and files
type(cfg_t) :: cfg
type(traj_t) :: traj
integer :: fu_stats
logical :: lvdw_stats, limud, has_rng, lexists
! run
type(c_ptr) :: ctx
type(bd_opts_t) :: opts
type(bd_runinfo_t) :: info
integer(c_int) :: rc
integer(c_int64_t) :: nts, nts_beg, nxt, used, seed, offset
integer(c_int), allocatable :: bonds(:,:), bond_styles(:), vdw_styles(:), mol(:,:)
real(rp), allocatable :: bond_params(:,:), vdw_params(:,:)
real(rp) :: box(3), st(30)
integer :: i, iarg, fu, nvdw, nbt, cntr = 0
real(rp), allocatable, target :: xi(:)
integer(c_int64_t) :: xi_pos = 0
integer(8) :: t0, t1, rate

fn_control = 'control.txt'; job_tag = ''
do iarg = 1, command_argument_count()
    call get_command_argument(iarg, arg)
    if (index(arg, 'fn_control=') == 1) fn_control = arg(12:)
    if (index(arg, 'job_tag=') == 1) job_tag = arg(9:)
end do

call read_control(trim(fn_control))

! ---- configuration: a new run reads the .cfg, a revived run the dump
has_rng = .false.
if (lrevive) then
    call read_dump(trim(fn_revive)//trim(job_tag), nts_beg, cfg, has_rng, seed, offset)
    print '(3a)', 'read revive file `', trim(fn_revive)//trim(job_tag), '`'
else
    call read_config(trim(fn_cfg)//trim(job_tag), cfg)
    nts_beg = 0
end if
call check_supported()

select case (trim(bdintg))
case ('EM'); opts%integrator = BD_EM
case ('PC'); opts%integrator = BD_PC
case ('SE'); opts%integrator = BD_SE
case default
    print '(a)', 'bdrun: bdintg must be EM, SE or PC'
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
opts%se_nlmxitr = se_nlmxitr
opts%se_kdmax = se_kdmax
opts%se_ftol = se_tol(1)
opts%se_stptol = se_tol(2)
opts%flow_style = cfg%flow_style
if (cfg%flow_style /= 0) opts%flow_rate = cfg%flow_params(1)

! ---- topology in the device's arrays
nbt = max(cfg%num_bond_types, 1)
allocate(bonds(3, max(cfg%num_bonds, 1)), bond_styles(nbt), bond_params(mxparam, nbt))
bonds = 0; bond_styles = 0; bond_params = 0
if (cfg%num_bonds > 0) bonds(:, 1:cfg%num_bonds) = cfg%bonds
if (cfg%num_bond_types > 0) then
    bond_styles(1:cfg%num_bond_types) = cfg%bond_styles
    bond_params(:, 1:cfg%num_bond_types) = cfg%bond_params
end if
call vdw_table()
limud = (cfg%num_tethers == 0) .and. (cfg%imcon == 0)
box = [cfg%basis(1,1), cfg%basis(2,2), cfg%basis(3,3)]

ctx = bd_ctx_create(int(cfg%num_atoms, c_int), opts)
if (.not. c_associated(ctx)) call fail(-1)
rc = bd_set_box(ctx, int(cfg%imcon, c_int), box); if (rc /= BD_OK) call fail(rc)
rc = bd_set_topology(ctx, int(cfg%num_bonds, c_int), bonds, int(cfg%num_bond_types, c_int), bond_styles, &
        bond_params, int(max(cfg%num_atom_types, 1), c_int), cfg%atoms, int(nvdw, c_int), vdw_styles, &
        vdw_params, merge(1, 0, lvdw), int(excluded_atoms, c_int), pt_code(mth_ptgen), rcutoff, tskin)
if (rc /= BD_OK) call fail(rc)
call set_terms()
rc = bd_set_coordinates(ctx, cfg%coordinates); if (rc /= BD_OK) call fail(rc)
if (lrevive) then
    ! mid-run coordinates: no second centring (bd_resume_run)
    rc = bd_resume_run(ctx, cfg%molecule_com); if (rc /= BD_OK) call fail(rc)
end if

! ---- random numbers (setup_m.f90:99-105)
if (.not. has_rng) then
    offset = 0
    if (read_seed) then
        open(newunit=fu, file='random_seed.txt'//trim(job_tag), action='read', status='old')
        read(fu, *) seed
        close(fu)
    else
        seed = seed_from(job_tag)
    end if
end if
if (write_seed) then
    open(newunit=fu, file='random_seed.txt'//trim(job_tag), action='write', status='replace')
    write(fu, *) seed
    close(fu)
end if
rc = bd_set_rng(ctx, seed, offset); if (rc /= BD_OK) call fail(rc)

! ---- an injected Gaussian stream in place of the device generator
if (len_trim(fn_xi) > 0) call read_xi(trim(fn_xi))
if (trim(device_path) == 'percall' .and. (.not. allocated(xi) .or. nts_samp > 0 .or. opts%integrator /= BD_EM)) then
    print '(a)', 'bdrun: device_path percall needs fn_xi, bdintg EM and nts_samp 0'
    stop 2
end if

! ---- statistics (stats_init) and trajectory (setup_m.f90:62, 69-97)
lvdw_stats = lvdw .and. (cfg%num_vdw_types > 0)
call stats_open(trim(fn_stats)//trim(job_tag), lrevive, cfg, lvdw_stats, fu_stats)
allocate(mol(2, max(cfg%num_molecules, 1)))
mol = 0
if (cfg%num_molecules > 0) mol(:, 1:cfg%num_molecules) = cfg%molecules(2:3, :)
rc = bd_stats_enable(ctx, int(stats_binsize(1), c_int), int(stats_binsize(2), c_int), &
        int(merge(cfg%num_molecules, 0, cfg%imcon /= 0), c_int), mol)
if (rc /= BD_OK) call fail(rc)
if (write_traj) then
    inquire(file=trim(fn_traj)//trim(job_tag), exist=lexists)
    if (lrevive .and. lexists) then
        call traj_open(trim(fn_traj)//trim(job_tag), traj)
    else
        call traj_create(trim(fn_traj)//trim(job_tag), cfg%num_atoms, traj)
    end if
end if

! ---- bds_run (bd_solver_m.f90:128-212)
call system_clock(t0, rate)
nts = nts_beg
rc = bd_run(ctx, 0_c_int, c_null_ptr, 0_c_long_long, used)     ! centring of an isolated molecule
if (rc /= BD_OK) call fail(rc)
print '(a,i0)', 'nts = ', nts
call dump(nts)
do
    if (nts > nts_sim) exit
    nxt = next_event(nts)
    if (trim(device_path) == 'percall') then
        call steps_percall(nxt - nts)
    else if (allocated(xi)) then
        if (xi_pos >= size(xi, kind=c_int64_t)) then
            print '(a)', 'bdrun: the fn_xi stream is exhausted'
            stop 1
        end if
        rc = bd_run(ctx, int(nxt - nts, c_int), c_loc(xi(xi_pos + 1)), size(xi, kind=c_int64_t) - xi_pos, used)
        xi_pos = xi_pos + used
    else
        rc = bd_run(ctx, int(nxt - nts, c_int), c_null_ptr, 0_c_long_long, used)
    end if
    if (rc /= BD_OK .and. rc /= BD_WARN_NOTCONV) then
        ! bds_run returns without the final dump (bd_solver_m.f90:171)
        print '(a,i0)', 'bdrun: failed in the steps after nts = ', nts
        call fail(rc)
    end if
    nts = nxt
    if (every(nts, nts_log)) print '(a,i0)', 'nts = ', nts
    if (every(nts, nts_dump)) call dump(nts)
    if (every(nts, nts_samp)) then
        rc = bd_stats_get(ctx, st); if (rc /= BD_OK) call fail(rc)
        call stats_write_line(fu_stats, nts, cfg, lvdw_stats, st)
        if (write_traj) then
            rc = bd_get_coordinates(ctx, cfg%coordinates); if (rc /= BD_OK) call fail(rc)
            call traj_append(traj, nts, cfg%coordinates)
        end if
    end if
end do
call dump(nts)
call system_clock(t1)
rc = bd_get_runinfo(ctx, info)
print '(a,es12.4)', 'execution time(s) ', real(t1 - t0, rp) / real(rate, rp)
print '(a,i0,a,i0)', 'mobility updates ', info%n_mob_updates, '  Lanczos iterations ', info%total_lanczos_iters
open(newunit=fu, file=trim(fn_out)//trim(job_tag), status='replace', action='write')
write(fu, '(i0)') cfg%num_atoms
write(fu, '(a,i0)') 'nts = ', nts
do i = 1, cfg%num_atoms
    write(fu, '(a,3es24.16)') 'M ', cfg%coordinates(:, i)
end do
close(fu)
close(fu_stats)
if (write_traj) call traj_close(traj)
call bd_ctx_destroy(ctx)

contains

subroutine fail(code)
    integer(c_int), intent(in) :: code
    print '(a,i0,2a)', 'bdrun: error ', code, ': ', bd_error_message()
    stop 1
end subroutine

logical function every(n, k)
    integer(c_int64_t), intent(in) :: n, k
    every = .false.
    if (k > 0) every = (mod(n, k) == 0)
end function

!> The first nts after n at which the loop logs, dumps or samples, or at
!> which it ends (nts_sim + 1); at most a million steps per bd_run call.
function next_event(n) result(t)
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t) :: t
    t = min(nts_sim + 1, n + 1000000)
    if (nts_log > 0) t = min(t, (n / nts_log + 1) * nts_log)
    if (nts_dump > 0) t = min(t, (n / nts_dump + 1) * nts_dump)
    if (nts_samp > 0) t = min(t, (n / nts_samp + 1) * nts_samp)
end function

!> write_dump at nts of the device state: coordinates, molecule_com, and
!> the generator's position.
subroutine dump(n)
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t) :: sd, off
    rc = bd_get_coordinates(ctx, cfg%coordinates); if (rc /= BD_OK) call fail(rc)
    rc = bd_get_runinfo(ctx, info); if (rc /= BD_OK) call fail(rc)
    if (limud .and. trim(device_path) /= 'percall') cfg%molecule_com = info%molecule_com
    rc = bd_get_rng(ctx, sd, off); if (rc /= BD_OK) call fail(rc)
    call write_dump(trim(fn_revive)//trim(job_tag), n, cfg, sd, off)
end subroutine

!> The injected Gaussian stream: raw float64 values.
subroutine read_xi(fn)
    character(len=*), intent(in) :: fn
    integer :: fx
    integer(c_int64_t) :: nbytes
    open(newunit=fx, file=fn, access='stream', form='unformatted', action='read', status='old')
    inquire(unit=fx, size=nbytes)
    allocate(xi(nbytes / 8))
    read(fx) xi
    close(fx)
end subroutine

!> k steps of integrate_em and the bds_run tail (bd_solver_m.f90:167-238,
!> 355-454) composed on the host from the entry points that replace its
!> call sites: bd_rpy_assemble (calc_rpy_tensor), bd_brownian
!> (brn_calc_bdw), bd_forces (ia_calc_forces) and bd_mob_symv (the dsymv of
!> calc_drift); the flow field, the update, the centre-of-mass removal and
!> the wrap are done here, as the reference does them.
subroutine steps_percall(k)
    integer(c_int64_t), intent(in) :: k
    real(rp), allocatable, save :: crd(:,:), frc(:,:), drift(:), dif(:,:), bdw(:,:), tmp(:)
    real(rp) :: en3(3), st9(9), bl3(3), com(3), sq2dt
    integer(c_int) :: lconv, iters
    integer :: n, s, col, ia, d
    integer(c_int64_t) :: it
    n = 3 * cfg%num_atoms
    s = merge(max(nts_mobsam, 1), 1, lhdia)
    sq2dt = sqrt(2.0_rp * tim_stp)
    if (.not. allocated(crd)) then
        allocate(crd(3, cfg%num_atoms), frc(3, cfg%num_atoms), drift(n), dif(n, s), bdw(n, s), tmp(n))
        dif = 0
    end if
    rc = bd_get_coordinates(ctx, crd); if (rc /= BD_OK) call fail(rc)
    do it = 1, k
        ! calc_diffusion
        if (.not. lhdia) then
            dif(:, 1) = sq2dt * xi(xi_pos + 1:xi_pos + n)
            xi_pos = xi_pos + n
            col = 1
        else
            if (cntr == 0) then
                rc = bd_set_coordinates(ctx, crd); if (rc /= BD_OK) call fail(rc)
                rc = bd_rpy_assemble(ctx); if (rc /= BD_OK) call fail(rc)
                rc = bd_brownian(ctx, xi(xi_pos + 1:), int(s, c_int), opts%mob_fctr, bdw, lconv, iters)
                if (rc /= BD_OK) call fail(rc)
                xi_pos = xi_pos + int(n, c_int64_t) * s
                dif = sq2dt * bdw
            else if (.not. keep_reuse_noise) then
                dif = 0
            end if
            cntr = cntr + 1
            col = cntr
        end if
        ! calc_drift
        rc = bd_set_coordinates(ctx, crd); if (rc /= BD_OK) call fail(rc)
        rc = bd_forces(ctx, frc, en3, st9, bl3); if (rc /= BD_OK) call fail(rc)
        drift = 0
        do ia = 1, cfg%num_atoms
            select case (cfg%flow_style)
            case (1)
                drift(3*ia-2) = opts%flow_rate * crd(2, ia)
            case (2)
                drift(3*ia-2) = opts%flow_rate * crd(1, ia)
                drift(3*ia-1) = -opts%flow_rate * crd(2, ia)
            case (3)
                drift(3*ia-2) = opts%flow_rate * crd(1, ia)
                drift(3*ia-1) = -0.5_rp * opts%flow_rate * crd(2, ia)
                drift(3*ia) = -0.5_rp * opts%flow_rate * crd(3, ia)
            end select
        end do
        if (lhdia) then
            rc = bd_mob_symv(ctx, frc, 1_c_int, tmp); if (rc /= BD_OK) call fail(rc)
            drift = tmp + drift
        else
            drift = drift + reshape(frc, [n])
        end if
        drift = tim_stp * drift
        ! integrate_em
        crd = crd + reshape(drift, [3, cfg%num_atoms]) + reshape(dif(:, col), [3, cfg%num_atoms])
        if (lhdia .and. cntr == s) cntr = 0
        ! the bds_run tail: centre of mass (to_origin) or wrap
        if (limud) then
            com = sum(crd, 2) / cfg%num_atoms
            do ia = 1, cfg%num_atoms
                crd(:, ia) = crd(:, ia) - com
            end do
            cfg%molecule_com = cfg%molecule_com + com
        else if (cfg%imcon /= 0) then
            do ia = 1, cfg%num_atoms
                do d = 1, 3
                    crd(d, ia) = crd(d, ia) - box(d) * floor(crd(d, ia) / box(d))
                end do
            end do
        end if
    end do
    rc = bd_set_coordinates(ctx, crd); if (rc /= BD_OK) call fail(rc)
end subroutine

!> Parts of atmcfg_t the device step does not cover: the statistics of
!> branched molecules (compute_ic0's backbone and side-chain terms).
!> Dihedrals add no force (ia_dihedral_m.f90 has no styles).
subroutine check_supported()
    if (cfg%num_branches > 0) then
        print '(a)', 'bdrun: branched molecules (BRANCHES) are not supported on the device'
        stop 2
    end if
end subroutine

!> Charges, angles, tethers and external fields (ia_calc_forces,
!> interaction_m.f90:84-147).
subroutine set_terms()
    if (cfg%num_atoms > 0) then
        rc = bd_set_charges(ctx, cfg%charge); if (rc /= BD_OK) call fail(rc)
    end if
    if (cfg%num_angles > 0) then
        rc = bd_set_angles(ctx, int(cfg%num_angles, c_int), cfg%angles, int(cfg%num_angle_types, c_int), &
                cfg%angle_styles, cfg%angle_params)
        if (rc /= BD_OK) call fail(rc)
    end if
    if (cfg%num_tethers > 0) then
        rc = bd_set_tethers(ctx, int(cfg%num_tethers, c_int), cfg%tethers, cfg%tether_points, &
                int(cfg%num_tether_types, c_int), cfg%tether_styles, cfg%tether_params)
        if (rc /= BD_OK) call fail(rc)
    end if
    if (cfg%num_externals > 0) then
        rc = bd_set_externals(ctx, int(cfg%num_externals, c_int), cfg%external_styles, cfg%external_params)
        if (rc /= BD_OK) call fail(rc)
    end if
end subroutine

!> vdw_styles/params in the .cfg's row order: row typ serves the atom-type
!> pair that interaction_m.f90:182-186 maps to typ.
subroutine vdw_table()
    nvdw = cfg%num_vdw_types
    allocate(vdw_styles(max(nvdw, 1)), vdw_params(mxparam, max(nvdw, 1)))
    vdw_styles = 0; vdw_params = 0
    if (nvdw > 0) then
        vdw_styles(1:nvdw) = cfg%vdw_styles
        vdw_params(:, 1:nvdw) = cfg%vdw_params
    end if
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
!> keep_reuse_noise, fn_out, fn_xi (a raw float64 file: the Gaussian stream
!> brn_calc_dw would draw, in place of the device generator) and
!> device_path (run: bd_run; percall: integrate_em composed on the host from
!> the per-call-site entries, see steps_percall).
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
        case ('stats_binsize'); read(val, *) stats_binsize
        case ('se_nlmxitr'); read(val, *) se_nlmxitr
        case ('se_kdmax'); read(val, *) se_kdmax
        case ('se_tol'); read(val, *) se_tol
        case ('tim_stp'); read(val, *) tim_stp
        case ('nts_sim'); read(val, *) d; nts_sim = int(d, c_int64_t)
        case ('nts_mobsam'); read(val, *) d; nts_mobsam = int(d)
        case ('nts_dump'); read(val, *) d; nts_dump = int(d, c_int64_t)
        case ('nts_samp'); read(val, *) d; nts_samp = int(d, c_int64_t)
        case ('nts_log'); read(val, *) d; nts_log = int(d, c_int64_t)
        case ('lrevive'); read(val, *) lrevive
        case ('read_seed'); read(val, *) read_seed
        case ('write_seed'); read(val, *) write_seed
        case ('write_traj'); read(val, *) write_traj
        case ('fn_cfg'); fn_cfg = trim(val)
        case ('fn_revive'); fn_revive = trim(val)
        case ('fn_stats'); fn_stats = trim(val)
        case ('fn_traj'); fn_traj = trim(val)
        case ('fn_out'); fn_out = trim(val)
        case ('fn_xi'); fn_xi = trim(val)
        case ('device_path'); read(val, *) device_path
        case ('precision'); read(val, *) precision
        case ('lanc_reorth'); read(val, *) lanc_reorth
        case ('keep_reuse_noise'); read(val, *) keep_reuse_noise
        end select
    end do
    close(fu)
end subroutine

end program bdrun
