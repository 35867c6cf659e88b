!> iso_c_binding interfaces to libbdstep.so (include/bdstep.h).
!>
!> This is the Fortran side of the drop-in boundary: the calls made here
!> replace calc_rpy_tensor (src/bd_solver_m.f90:458-522), brn_calc_bdw
!> (src/brown_m.f90:147-181), ia_calc_forces (src/interaction_m.f90:84-147),
!> the dsymv drift product (src/bd_solver_m.f90:446) and the integrate_em
!> step loop (src/bd_solver_m.f90:167-238).
module bdstep_c
use iso_c_binding
implicit none

integer(c_int), parameter :: BD_OK = 0, BD_ERR_BOND = 1, BD_ERR_FACTOR = 2, &
    BD_WARN_NOTCONV = 3, BD_ERR_SOLVE = 4
integer(c_int), parameter :: BD_CHOL = 0, BD_LANC = 1, BD_CHOL_LT = 2
integer(c_int), parameter :: BD_FP64 = 0, BD_FP32 = 1
integer(c_int), parameter :: BD_PT_DIR = 0, BD_PT_VER = 1, BD_PT_CL = 2
integer(c_int), parameter :: BD_EM = 0, BD_PC = 1, BD_SE = 2
integer(c_int), parameter :: BD_MOB_AUTO = 0, BD_MOB_TILES = 1, BD_MOB_MATFREE = 2

type, bind(C) :: bd_opts_t
    integer(c_int) :: lhdia = 1
    integer(c_int) :: mob_fctr = BD_CHOL
    integer(c_int) :: nts_mobsam = 1
    integer(c_int) :: lanc_mxitr = 0
    real(c_double) :: lanc_tol = 0.0_c_double
    real(c_double) :: tim_stp = 0.0_c_double
    integer(c_int) :: precision = BD_FP64
    integer(c_int) :: keep_reuse_noise = 0
    integer(c_int) :: flow_style = 0
    real(c_double) :: flow_rate = 0.0_c_double
    integer(c_int) :: lanc_reorth = 0
    integer(c_int) :: integrator = 0
    integer(c_int) :: mob_storage = BD_MOB_AUTO
    integer(c_int) :: se_nlmxitr = 0
    integer(c_int) :: se_kdmax = 0
    real(c_double) :: se_ftol = 0.0_c_double
    real(c_double) :: se_stptol = 0.0_c_double
end type

type, bind(C) :: bd_runinfo_t
    integer(c_long_long) :: nts
    real(c_double) :: molecule_com(3)
    integer(c_int) :: last_lanczos_iters
    integer(c_int) :: total_lanczos_iters
    integer(c_int) :: n_mob_updates
    integer(c_int) :: n_not_converged
    integer(c_int) :: n_pair_rebuilds
end type

interface
    function bd_ctx_create(nbeads, opts) bind(C, name="bd_ctx_create") result(ctx)
        import :: c_int, c_ptr, bd_opts_t
        integer(c_int), value :: nbeads
        type(bd_opts_t), intent(in) :: opts
        type(c_ptr) :: ctx
    end function
    subroutine bd_ctx_destroy(ctx) bind(C, name="bd_ctx_destroy")
        import :: c_ptr
        type(c_ptr), value :: ctx
    end subroutine
    function bd_last_error() bind(C, name="bd_last_error") result(p)
        import :: c_ptr
        type(c_ptr) :: p
    end function
    function bd_set_box(ctx, imcon, box) bind(C, name="bd_set_box") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: imcon
        real(c_double), intent(in) :: box(3)
        integer(c_int) :: rc
    end function
    function bd_set_topology(ctx, nbonds, bonds, nbond_types, bond_styles, bond_params, &
            natom_types, atom_types, nvdw_types, vdw_styles, vdw_params, lvdw, &
            excluded_atoms, mth_ptgen, rcutoff, tskin) bind(C, name="bd_set_topology") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: nbonds, nbond_types, natom_types, nvdw_types
        integer(c_int), intent(in) :: bonds(*), bond_styles(*), atom_types(*), vdw_styles(*)
        real(c_double), intent(in) :: bond_params(*), vdw_params(*)
        integer(c_int), value :: lvdw, excluded_atoms, mth_ptgen
        real(c_double), value :: rcutoff, tskin
        integer(c_int) :: rc
    end function
    function bd_set_coordinates(ctx, crd) bind(C, name="bd_set_coordinates") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: crd(*)
        integer(c_int) :: rc
    end function
    function bd_get_coordinates(ctx, crd) bind(C, name="bd_get_coordinates") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(out) :: crd(*)
        integer(c_int) :: rc
    end function
    function bd_rpy_assemble(ctx) bind(C, name="bd_rpy_assemble") result(rc)
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx
        integer(c_int) :: rc
    end function
    function bd_brownian(ctx, xi, s, method, bdw, lconv, iters) bind(C, name="bd_brownian") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: xi(*)
        integer(c_int), value :: s, method
        real(c_double), intent(out) :: bdw(*)
        integer(c_int), intent(out) :: lconv, iters
        integer(c_int) :: rc
    end function
    function bd_forces(ctx, frc, energies, stress9, bndlen3) bind(C, name="bd_forces") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(out) :: frc(*), energies(3), stress9(9), bndlen3(3)
        integer(c_int) :: rc
    end function
    function bd_run(ctx, nsteps, xi, nxi, used) bind(C, name="bd_run") result(rc)
        import :: c_ptr, c_int, c_long_long
        type(c_ptr), value :: ctx
        integer(c_int), value :: nsteps
        type(c_ptr), value :: xi
        integer(c_long_long), value :: nxi
        integer(c_long_long), intent(out) :: used
        integer(c_int) :: rc
    end function
    function bd_set_rng(ctx, seed, offset) bind(C, name="bd_set_rng") result(rc)
        import :: c_ptr, c_int, c_int64_t
        type(c_ptr), value :: ctx
        integer(c_int64_t), value :: seed, offset
        integer(c_int) :: rc
    end function
    function bd_get_rng(ctx, seed, offset) bind(C, name="bd_get_rng") result(rc)
        import :: c_ptr, c_int, c_int64_t
        type(c_ptr), value :: ctx
        integer(c_int64_t), intent(out) :: seed, offset
        integer(c_int) :: rc
    end function
    function bd_resume_run(ctx, com) bind(C, name="bd_resume_run") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: com(3)
        integer(c_int) :: rc
    end function
    function bd_stats_enable(ctx, binsize, binsize_stress, nmolecules, molecules) &
            bind(C, name="bd_stats_enable") result(rc)
        import :: c_ptr, c_int
        type(c_ptr), value :: ctx
        integer(c_int), value :: binsize, binsize_stress, nmolecules
        integer(c_int), intent(in) :: molecules(*)
        integer(c_int) :: rc
    end function
    function bd_stats_get(ctx, out) bind(C, name="bd_stats_get") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(out) :: out(30)
        integer(c_int) :: rc
    end function
    function bd_mob_symv(ctx, x, s, y) bind(C, name="bd_mob_symv") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: x(*)
        integer(c_int), value :: s
        real(c_double), intent(out) :: y(*)
        integer(c_int) :: rc
    end function
    function bd_set_angles(ctx, nangles, angles, nangle_types, angle_styles, angle_params) &
            bind(C, name="bd_set_angles") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: nangles, nangle_types
        integer(c_int), intent(in) :: angles(*), angle_styles(*)
        real(c_double), intent(in) :: angle_params(*)
        integer(c_int) :: rc
    end function
    function bd_set_tethers(ctx, ntethers, tethers, tether_points, ntether_types, tether_styles, &
            tether_params) bind(C, name="bd_set_tethers") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: ntethers, ntether_types
        integer(c_int), intent(in) :: tethers(*), tether_styles(*)
        real(c_double), intent(in) :: tether_points(*), tether_params(*)
        integer(c_int) :: rc
    end function
    function bd_set_externals(ctx, nexternals, external_styles, external_params) &
            bind(C, name="bd_set_externals") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: nexternals
        integer(c_int), intent(in) :: external_styles(*)
        real(c_double), intent(in) :: external_params(*)
        integer(c_int) :: rc
    end function
    function bd_set_charges(ctx, charge) bind(C, name="bd_set_charges") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: charge(*)
        integer(c_int) :: rc
    end function
    function bd_get_runinfo(ctx, info) bind(C, name="bd_get_runinfo") result(rc)
        import :: c_ptr, c_int, bd_runinfo_t
        type(c_ptr), value :: ctx
        type(bd_runinfo_t), intent(out) :: info
        integer(c_int) :: rc
    end function
    function bd_rgsq(ctx, rgsq) bind(C, name="bd_rgsq") result(rc)
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(out) :: rgsq
        integer(c_int) :: rc
    end function
    function c_strlen(p) bind(C, name="strlen") result(n)
        import :: c_ptr, c_size_t
        type(c_ptr), value :: p
        integer(c_size_t) :: n
    end function
end interface

contains

!> Message of the last failing call, as a Fortran string.
function bd_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: n, i
    p = bd_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, cs, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
        msg(i:i) = cs(i)
    end do
end function

end module bdstep_c
