! ---------------------------------------------------------------------------
! mpm_hip -- ISO_C_BINDING interface to libmpm3d_hip.so (include/mpm3d_hip.h,
! include/mpm3d_problem.h).
!
! The derived types mirror the C structs field for field (bind(C)); every
! entry point returns an integer(c_int) status, 0 on success.
! ---------------------------------------------------------------------------
module mpm_hip
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: MPM_PREC_FP64 = 0, MPM_PREC_FP32 = 1, MPM_PREC_FP64_EXACT = 2
  integer(c_int), parameter :: MPM_USF = 0, MPM_USL = 1, MPM_MUSL = 2
  integer(c_int), parameter :: MPM_SHAPE_LINEAR = 0, MPM_SHAPE_GIMP = 1, &
                               MPM_SHAPE_BSPLINE = 2, MPM_SHAPE_SGMP = 3
  ! particle field indices (0-based rows of pstate, enum mpm_pfield)
  integer, parameter :: PF_X = 0, PF_Y = 1, PF_Z = 2, PF_VX = 3, PF_VY = 4, PF_VZ = 5, &
                        PF_FX = 6, PF_FY = 7, PF_FZ = 8, PF_MASS = 9, PF_VOL = 10, PF_SM = 11, &
                        PF_SEQV = 12, PF_SDXX = 13, PF_SDYY = 14, PF_SDZZ = 15, PF_SDYZ = 16, &
                        PF_SDXZ = 17, PF_SDXY = 18, PF_SIGY = 19, PF_EPEFF = 20, PF_IE = 21, &
                        PF_Q = 22, PF_CP = 23, PF_FP = 24, PF_XX = 33, PF_TEMP = 36, PF_DMG = 37, &
                        PF_LT = 38, PF_FAIL = 39, PF_SPRE = 40, PF_XO = 41, PF_VOL0 = 44, PF_PK = 45, &
                        PF_COUNT = 54

  type, bind(C) :: mpm_material_t
     integer(c_int32_t) :: mtype, eos_type
     real(c_double) :: density, young, poisson, yield0, tangmod, miu, bulk_modulus, wavespd
     real(c_double) :: q_fai, k_fai, q_psi, ten_f
     real(c_double) :: ceos(10)
     real(c_double) :: det_speed
     real(c_double) :: b_jc, n_jc, c_jc, m_jc, epso, epf, roomt, melt, spec_heat
  end type mpm_material_t

  type, bind(C) :: mpm_body_t
     integer(c_int32_t) :: mat, com, par_begin, par_end
     real(c_double) :: grav(3)
  end type mpm_body_t

  type, bind(C) :: mpm_problem_t
     integer(c_int32_t) :: nb_particle, nb_body, nb_mat, shape, scheme
     integer(c_int32_t) :: quasi_load, quasi_damp, gravity, update_v_by_fp
     integer(c_int32_t) :: fix(6)
     integer(c_int32_t) :: ncx, ncy, ncz, dr_damping
     real(c_double) :: span(6), dcell, cutoff, dt, dtscale, endtime
     real(c_double) :: quasi_load_time, damp_factor, bq1, bq2
     type(c_ptr) :: mats, bodies, pstate, pbody
     integer(c_int32_t) :: smooth_stress, tlmpm, contact, normbody, nb_component, reserved0
     real(c_double) :: fricfa
  end type mpm_problem_t

  interface
     function mpm_last_error() bind(C, name="mpm_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function mpm_last_error

     function mpm_deck_read(path, prob, errbuf, errlen) bind(C, name="mpm_deck_read") result(rc)
       import :: c_int, c_char, c_ptr
       character(kind=c_char), dimension(*), intent(in) :: path
       type(c_ptr), intent(out) :: prob
       character(kind=c_char), dimension(*), intent(out) :: errbuf
       integer(c_int), value :: errlen
       integer(c_int) :: rc
     end function mpm_deck_read

     subroutine mpm_problem_free(prob) bind(C, name="mpm_problem_free")
       import :: c_ptr
       type(c_ptr), value :: prob
     end subroutine mpm_problem_free

     function mpm_ctx_create(prob, precision, ctx) bind(C, name="mpm_ctx_create") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: prob
       integer(c_int), value :: precision
       type(c_ptr), intent(out) :: ctx
       integer(c_int) :: rc
     end function mpm_ctx_create

     function mpm_ctx_destroy(ctx) bind(C, name="mpm_ctx_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_ctx_destroy

     function mpm_begin_step(ctx) bind(C, name="mpm_begin_step") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_begin_step

     function mpm_grid_momentum_initial(ctx) bind(C, name="mpm_grid_momentum_initial") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_grid_momentum_initial

     function mpm_apply_boundary_conditions(ctx) bind(C, name="mpm_apply_boundary_conditions") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_apply_boundary_conditions

     function mpm_particle_stress_update(ctx) bind(C, name="mpm_particle_stress_update") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_particle_stress_update

     function mpm_smooth_stress_by_grid(ctx) bind(C, name="mpm_smooth_stress_by_grid") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_smooth_stress_by_grid

     function mpm_lagr_nod_contact(ctx) bind(C, name="mpm_lagr_nod_contact") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_lagr_nod_contact

     function mpm_get_contact_force(ctx, out3) bind(C, name="mpm_get_contact_force") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out3(3)
       integer(c_int) :: rc
     end function mpm_get_contact_force

     function mpm_grid_momentum_update(ctx) bind(C, name="mpm_grid_momentum_update") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_grid_momentum_update

     function mpm_integrate_momentum(ctx) bind(C, name="mpm_integrate_momentum") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_integrate_momentum

     function mpm_particle_position_update(ctx) bind(C, name="mpm_particle_position_update") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_particle_position_update

     function mpm_grid_momentum_musl(ctx) bind(C, name="mpm_grid_momentum_musl") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_grid_momentum_musl

     function mpm_calc_energy(ctx, ek, ei, mom) bind(C, name="mpm_calc_energy") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ek, ei
       real(c_double), intent(out) :: mom(3)
       integer(c_int) :: rc
     end function mpm_calc_energy

     function mpm_update_dt(ctx, dt) bind(C, name="mpm_update_dt") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt
       integer(c_int) :: rc
     end function mpm_update_dt

     function mpm_step(ctx, nsteps) bind(C, name="mpm_step") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
       integer(c_int) :: rc
     end function mpm_step

     function mpm_download_particles(ctx, pstate) bind(C, name="mpm_download_particles") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: pstate(*)
       integer(c_int) :: rc
     end function mpm_download_particles

     function mpm_download_particle_subset(ctx, k, ids, out) bind(C, name="mpm_download_particle_subset") &
          result(rc)
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: k
       integer(c_int32_t), intent(in) :: ids(*)
       real(c_double), intent(out) :: out(*)
       integer(c_int) :: rc
     end function mpm_download_particle_subset

     function mpm_download_icell(ctx, icell) bind(C, name="mpm_download_icell") result(rc)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: icell(*)
       integer(c_int) :: rc
     end function mpm_download_icell

     function mpm_upload_particles(ctx, pstate) bind(C, name="mpm_upload_particles") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: pstate(*)
       integer(c_int) :: rc
     end function mpm_upload_particles

     function mpm_get_scalars(ctx, out8) bind(C, name="mpm_get_scalars") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out8(8)
       integer(c_int) :: rc
     end function mpm_get_scalars

     function mpm_sync(ctx) bind(C, name="mpm_sync") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mpm_sync

     function c_strlen(s) bind(C, name="strlen") result(n)
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: n
     end function c_strlen
  end interface

contains

  ! stop with the library's message when a call fails (the reference stops on
  ! fatal input/model errors the same way)
  subroutine mpm_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(len=512) :: msg
    if (rc /= 0) then
       call mpm_error_text(msg)
       write(*, '(a,a,a,i0,a,a)') '*** Error *** ', where, ' failed (', rc, '): ', trim(msg)
       stop 'mpm_hip: device engine error'
    end if
  end subroutine mpm_check

  subroutine mpm_error_text(msg)
    character(len=*), intent(out) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    msg = ' '
    p = mpm_last_error()
    if (.not. c_associated(p)) return
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    do i = 1, min(n, len(msg))
       msg(i:i) = chars(i)
    end do
  end subroutine mpm_error_text

end module mpm_hip
