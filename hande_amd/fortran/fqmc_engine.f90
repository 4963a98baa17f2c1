! ISO_C_BINDING interface to the MI355X FCIQMC engine (include/fqmc.h).
!
! This module is what the reference's do_fciqmc (src/fciqmc.f90:12) would
! `use` to hand the body of its Monte Carlo cycle (src/fciqmc.f90:307-399) to
! the GPU.  The derived types are bind(c) mirrors of the C structs; every
! function returns the engine's integer status (FQ_OK = 0).
module fqmc_engine

    use, intrinsic :: iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_ptr, c_null_ptr

    implicit none

    integer(c_int32_t), parameter :: FQ_OK = 0
    integer(c_int32_t), parameter :: FQ_SYS_HUB_K = 1, FQ_SYS_READ_IN = 2
    integer(c_int32_t), parameter :: FQ_EXCIT_RENORM = 0, FQ_EXCIT_NO_RENORM = 1, FQ_EXCIT_HEAT_BATH = 2
    integer(c_int32_t), parameter :: FQ_EXCIT_LEARNED = 3
    integer(c_int32_t), parameter :: FQ_RNG_PHILOX = 0, FQ_RNG_REPLAY = 1
    integer, parameter :: FQ_MAX_SPACES = 2

    ! Read-only system tables (pointers to host arrays owned by the caller).
    type, bind(c) :: fq_system
        integer(c_int32_t) :: system, nbasis, nel, nalpha, nbeta
        integer(c_int32_t) :: nvirt, nvirt_alpha, nvirt_beta, nwords
        type(c_ptr) :: f0 = c_null_ptr
        real(c_double) :: H00
        type(c_ptr) :: sp_eigv = c_null_ptr
        integer(c_int32_t) :: nsym
        type(c_ptr) :: sym_table = c_null_ptr, inv_sym = c_null_ptr
        real(c_double) :: coulomb_k
        integer(c_int32_t) :: norb, nsym_pg
        type(c_ptr) :: orb_sym = c_null_ptr, nbasis_sym_spin = c_null_ptr
        integer(c_int32_t) :: max_sym_spin
        type(c_ptr) :: sym_spin_basis_fns = c_null_ptr
        real(c_double) :: ecore
        type(c_ptr) :: h1 = c_null_ptr, eri = c_null_ptr
        integer(c_int64_t) :: n_eri
        real(c_double) :: pattempt_single, pattempt_double
    end type fq_system

    type, bind(c) :: fq_params
        integer(c_int32_t) :: nspaces, initiator
        real(c_double) :: initiator_pop
        integer(c_int32_t) :: quadrature_initiator, real_amplitudes, pop_bit_shift
        integer(c_int64_t) :: spawn_cutoff
        integer(c_int32_t) :: excit_gen, rng_mode
        integer(c_int64_t) :: philox_seed
        integer(c_int64_t) :: main_capacity, spawn_capacity
        integer(c_int32_t) :: device
        real(c_double) :: proposal_mix
        integer(c_int32_t) :: proposal_min_attempts
    end type fq_params

    type, bind(c) :: fq_cycle_out
        real(c_double) :: proj_energy(FQ_MAX_SPACES), D0_population(FQ_MAX_SPACES), nparticles(FQ_MAX_SPACES)
        integer(c_int64_t) :: nstates, nspawn_events, ndeath, nattempts, rng_consumed
        integer(c_int32_t) :: spawn_overflow, main_overflow
    end type fq_cycle_out

    interface

        function fq_create(sys, params, engine) result(rc) bind(c, name='fq_create')
            import :: fq_system, fq_params, c_ptr, c_int32_t
            type(fq_system), intent(in) :: sys
            type(fq_params), intent(in) :: params
            type(c_ptr), intent(out) :: engine
            integer(c_int32_t) :: rc
        end function fq_create

        function fq_load(engine, states, pops, dat, n) result(rc) bind(c, name='fq_load')
            import :: c_ptr, c_int64_t, c_double, c_int32_t
            type(c_ptr), value :: engine
            integer(c_int64_t), intent(in) :: states(*)   ! psip_list%states(:,1:n) with DET_SIZE=64
            integer(c_int64_t), intent(in) :: pops(*)     ! psip_list%pops(:,1:n) with POP_SIZE=64
            real(c_double), intent(in) :: dat(*)          ! psip_list%dat(1,1:n)
            integer(c_int64_t), value :: n
            integer(c_int32_t) :: rc
        end function fq_load

        function fq_push_rng(engine, draws, n) result(rc) bind(c, name='fq_push_rng')
            import :: c_ptr, c_int64_t, c_double, c_int32_t
            type(c_ptr), value :: engine
            real(c_double), intent(in) :: draws(*)
            integer(c_int64_t), value :: n
            integer(c_int32_t) :: rc
        end function fq_push_rng

        function fq_rng_available(engine, n) result(rc) bind(c, name='fq_rng_available')
            import :: c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: engine
            integer(c_int64_t), intent(out) :: n
            integer(c_int32_t) :: rc
        end function fq_rng_available

        function fq_cycle(engine, shift, eproj_old, tau, out) result(rc) bind(c, name='fq_cycle')
            import :: c_ptr, c_double, c_int32_t, fq_cycle_out
            type(c_ptr), value :: engine
            real(c_double), intent(in) :: shift(*), eproj_old(*)
            real(c_double), value :: tau
            type(fq_cycle_out), intent(out) :: out
            integer(c_int32_t) :: rc
        end function fq_cycle

        function fq_annihilate(engine, sstates, spops, sflags, ns, out) result(rc) bind(c, name='fq_annihilate')
            import :: c_ptr, c_int64_t, c_int32_t, fq_cycle_out
            type(c_ptr), value :: engine
            integer(c_int64_t), intent(in) :: sstates(*), spops(*), sflags(*)
            integer(c_int64_t), value :: ns
            type(fq_cycle_out), intent(out) :: out
            integer(c_int32_t) :: rc
        end function fq_annihilate

        function fq_nstates(engine, n) result(rc) bind(c, name='fq_nstates')
            import :: c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: engine
            integer(c_int64_t), intent(out) :: n
            integer(c_int32_t) :: rc
        end function fq_nstates

        function fq_download(engine, states, pops, dat, capacity, n) result(rc) bind(c, name='fq_download')
            import :: c_ptr, c_int64_t, c_double, c_int32_t
            type(c_ptr), value :: engine
            integer(c_int64_t), intent(out) :: states(*), pops(*)
            real(c_double), intent(out) :: dat(*)
            integer(c_int64_t), value :: capacity
            integer(c_int64_t), intent(out) :: n
            integer(c_int32_t) :: rc
        end function fq_download

        ! init_semi_stoch_t for space 'high' (separate = 1: separate annihilation)
        function fq_semi_stoch_init(engine, target_size, separate) result(rc) bind(c, name='fq_semi_stoch_init')
            import :: c_ptr, c_int64_t, c_int32_t
            type(c_ptr), value :: engine
            integer(c_int64_t), value :: target_size
            integer(c_int32_t), value :: separate
            integer(c_int32_t) :: rc
        end function fq_semi_stoch_init

        ! online training of the learned proposal (excit_gen FQ_EXCIT_LEARNED):
        ! no reference counterpart; objective 0 cross-entropy, 1 spawning variance
        function fq_mlp_online(engine, every, steps, batch, lr, objective) result(rc) bind(c, name='fq_mlp_online')
            import :: c_ptr, c_int64_t, c_int32_t, c_double
            type(c_ptr), value :: engine
            integer(c_int32_t), value :: every, steps
            integer(c_int64_t), value :: batch
            real(c_double), value :: lr
            integer(c_int32_t), value :: objective
            integer(c_int32_t) :: rc
        end function fq_mlp_online

        subroutine fq_destroy(engine) bind(c, name='fq_destroy')
            import :: c_ptr
            type(c_ptr), value :: engine
        end subroutine fq_destroy

    end interface

end module fqmc_engine
