! ---------------------------------------------------------------------------
! mpm3d_host -- Fortran host for the MI355X engine.
!
! Reads a .mpm deck with the library's reader (the reference's InputPara
! semantics), uploads it once, and runs the explicit updated-Lagrangian loop
! through the per-phase entry points in the call order of the reference
! driver (MPM3D.f90:69-160), including its end-of-run test and its
! "Time per particle cycle" report (MPM3D.f90:267,277).
!
! usage: mpm3d_host <deck.mpm> [nsteps] [fp32|fp64] [dumpfile]
!   nsteps    stop after this many steps instead of at EndTime (0: EndTime)
!   dumpfile  final particle state, PF_COUNT x nb_particle doubles (stream)
! ---------------------------------------------------------------------------
program mpm3d_host
  use, intrinsic :: iso_c_binding
  use mpm_hip
  implicit none

  character(len=1024) :: deck, arg, dumpfile
  character(kind=c_char) :: errbuf(512)
  type(c_ptr) :: pprob, ctx
  type(mpm_problem_t), pointer :: prob
  integer :: nsteps, istep, k, u, n
  integer(c_int) :: prec, rc
  integer(8) :: c0, c1, crate
  real(c_double) :: ek, ei, mom(3), dt, sc(8), t_loop
  real(c_double), allocatable :: ps(:)
  logical :: usf, musl, jc
  type(mpm_material_t), pointer :: mats(:)
  integer :: m

  if (command_argument_count() < 1) then
     print *, 'usage: mpm3d_host <deck.mpm> [nsteps] [fp32|fp64|fp64exact] [dumpfile]'
     stop 2
  end if
  call get_command_argument(1, deck)
  nsteps = 0
  if (command_argument_count() >= 2) then
     call get_command_argument(2, arg)
     read(arg, *) nsteps
  end if
  prec = MPM_PREC_FP64
  if (command_argument_count() >= 3) then
     call get_command_argument(3, arg)
     if (trim(arg) == 'fp32') prec = MPM_PREC_FP32
     if (trim(arg) == 'fp64exact') prec = MPM_PREC_FP64_EXACT
  end if
  dumpfile = ' '
  if (command_argument_count() >= 4) call get_command_argument(4, dumpfile)

  rc = mpm_deck_read(trim(deck)//c_null_char, pprob, errbuf, 512)
  if (rc /= 0) then
     write(*, '(a)') '*** Error *** reading deck: '//transfer(errbuf(1:200), repeat(' ', 200))
     stop 'mpm3d_host: bad deck'
  end if
  call c_f_pointer(pprob, prob)
  n = prob%nb_particle
  ! contact, total-Lagrangian and Johnson-Cook (john, jcf) problems run in
  ! the reference-order FP64 mode only
  call c_f_pointer(prob%mats, mats, [prob%nb_mat])
  jc = .false.
  do m = 1, prob%nb_mat
     jc = jc .or. mats(m)%mtype == 4 .or. mats(m)%mtype == 7
  end do
  if ((prob%contact /= 0 .or. prob%tlmpm /= 0 .or. jc) .and. prec /= MPM_PREC_FP64_EXACT) then
     write(*, '(a)') ' mpm3d_host: this deck runs in the reference-order FP64 mode (fp64exact)'
     prec = MPM_PREC_FP64_EXACT
  end if
  usf = prob%scheme == MPM_USF
  musl = prob%scheme == MPM_MUSL
  call mpm_check(mpm_ctx_create(pprob, prec, ctx), 'mpm_ctx_create')
  write(*, '(a,i12,a,es10.3)') ' particles = ', n, '   DT = ', prob%dt

  call system_clock(c0, crate)
  istep = 0
  do
     istep = istep + 1
     call mpm_check(mpm_begin_step(ctx), 'begin_step')                           ! MPM3D.f90:72-74
     call mpm_check(mpm_grid_momentum_initial(ctx), 'GridMomentumInitial')      ! :77
     if (usf) then
        call mpm_check(mpm_apply_boundary_conditions(ctx), 'ApplyBoundaryConditions')
        call mpm_check(mpm_particle_stress_update(ctx), 'ParticleStressUpdate')
     end if
     call mpm_check(mpm_grid_momentum_update(ctx), 'GridMomentumUpdate')        ! :87
     call mpm_check(mpm_integrate_momentum(ctx), 'IntegrateMomentum')           ! :90
     if (prob%contact == 1) call mpm_check(mpm_lagr_nod_contact(ctx), 'Lagr_NodContact')   ! :94-96
     call mpm_check(mpm_particle_position_update(ctx), 'ParticlePositionUpdate') ! :99
     if (musl) then
        call mpm_check(mpm_grid_momentum_musl(ctx), 'GridMomentumMUSL')
        call mpm_check(mpm_apply_boundary_conditions(ctx), 'ApplyBoundaryConditions')
     end if
     if (.not. usf) call mpm_check(mpm_particle_stress_update(ctx), 'ParticleStressUpdate')
     call mpm_check(mpm_calc_energy(ctx, ek, ei, mom), 'calcEnergy')            ! :112
     call mpm_check(mpm_update_dt(ctx, dt), 'UpdateDT')                         ! :113
     if (prob%smooth_stress /= 0) &
        call mpm_check(mpm_smooth_stress_by_grid(ctx), 'SmoothStressbyGrid')     ! :114
     call mpm_check(mpm_get_scalars(ctx, sc), 'get_scalars')
     if (nsteps > 0) then
        if (istep >= nsteps) exit
     else if (prob%quasi_load /= 0) then                                         ! :152-158
        if ((sc(2) > prob%endtime .and. ek / (ek + ei) < 0.005d0) .and. sc(2) > 100d0) exit
     else
        if (sc(2) > prob%endtime) exit
     end if
  end do
  call mpm_check(mpm_sync(ctx), 'sync')
  call system_clock(c1)
  t_loop = dble(c1 - c0) / dble(crate)

  write(*, '(a,i8,a,es12.4,a,es12.4,a,es12.4)') ' Step=', istep, '  T=', sc(2), '  K.E.=', ek, &
       '  T.E.=', ek + ei
  write(*, '(a,f12.4,a)') '** Time per particle cycle: ', t_loop / n / istep * 1.0d9, ' nanoseconds'

  if (len_trim(dumpfile) > 0) then
     allocate(ps(PF_COUNT * n))
     call mpm_check(mpm_download_particles(ctx, ps), 'download')
     open(newunit=u, file=trim(dumpfile), access='stream', form='unformatted', status='replace')
     write(u) int(n, 4), int(istep, 4), sc
     write(u) ps
     close(u)
  end if
  call mpm_check(mpm_ctx_destroy(ctx), 'destroy')
  call mpm_problem_free(pprob)
  k = 0
end program mpm3d_host
