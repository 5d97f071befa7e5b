! ---------------------------------------------------------------------------
! update_step_hip -- drop-in replacement of the reference's update_step.f90.
!
! Defines the same argument-less external subroutines the reference driver
! calls (MPM3D.f90:77-113): GridMomentumInitial, ApplyBoundaryConditions,
! ParticleStressUpdate, GridMomentumUpdate, IntegrateMomentum,
! ParticlePositionUpdate, GridMomentumMUSL, UpdateDT.  Linked in place of
! update_step.o, they run the step on the MI355X through libmpm3d_hip.so.
!
! State: on the first call the module globals built by InputPara
! (particle_list, body_list, mat_list, grid description, flags, DT) are
! packed into an mpm_problem and uploaded; from then on the device owns the
! state.  UpdateDT, the last routine of a step, returns the device DT into
! DT and refreshes what the driver reads next: the kinetic and internal
! energies and momentum (replacing the host sums calcEnergy just made from
! stale particles), the curve particles of OutCurve (a gather of those few
! particles on the device), and, on the steps where the driver writes its
! TecPlot / ParaView output (its own schedule, plt += OutTime, mirrored
! here), the whole particle state and the cell indices.
!
! Environment: MPM3D_PRECISION=fp32 selects the FP32-bulk path, =fp64exact
! the reference-order FP64 path (bit-identical to the reference);
! MPM3D_DUMP=<file> writes the particle state after every step (stream,
! n, istep, 8 scalars, PF_COUNT x n doubles).
!
! SmoothStressbyGrid, which the driver calls after UpdateDT, smooths on the
! device and, on output steps, refreshes the particles' Spre for the
! ParaView writer.  Lagr_NodContact runs the contact pass on the device;
! contact decks run in the reference-order FP64 mode (the only mode with
! grid components), and tot_cont_for is refreshed for OutCurve.
!
! The total-Lagrangian routines (TLmp decks, MPM3D.f90:162-250) map to the
! same entry points: the context knows the problem is total-Lagrangian
! (GridMassInitial only creates it: the device takes the initial mass in the
! first step), and such decks run in the reference-order FP64 mode.
!
! Decks with the Johnson-Cook models john and jcf also run in that mode: a
! particle that does not yield heats (and jcf damages) with the plastic
! strain increment of the last particle before it in the reference's order
! that did, which that mode keeps.
!
! Compiles against the reference's module files (ParticleData, GridData,
! MaterialData).
! ---------------------------------------------------------------------------
module update_step_hip_state
  use, intrinsic :: iso_c_binding
  use mpm_hip
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  type(mpm_problem_t), target, save :: prob
  type(mpm_material_t), allocatable, target, save :: mats(:)
  type(mpm_body_t), allocatable, target, save :: bodies(:)
  real(c_double), allocatable, target, save :: ps(:)
  integer(c_int32_t), allocatable, target, save :: pbody(:)
  integer, save :: nsteps_done = 0
  ! the driver's output schedule (MPM3D.f90:35, 47, 125-126): a default REAL
  real, save :: plt_mirror = 0.0
  logical, save :: plt_started = .false.
  real(c_double), allocatable, target, save :: sub(:)
  integer(c_int32_t), allocatable, target, save :: icell_buf(:)
  ! the last UpdateDT refreshed the whole particle view (an output step)
  logical, save :: full_view = .false.

contains

  subroutine ensure_context()
    use ParticleData
    use GridData
    use MaterialData
    integer :: i, m, b, n, j
    integer(c_int) :: prec
    character(len=16) :: env
    real(8) :: dt0, cp0
    logical :: jc
    if (c_associated(ctx)) return
    n = nb_particle
    prob%nb_particle = n
    prob%nb_body = nb_body
    prob%nb_mat = nb_mat
    if (SGMP) then
       prob%shape = MPM_SHAPE_SGMP
    else if (GIMP) then
       prob%shape = MPM_SHAPE_GIMP
    else if (Bspline) then
       prob%shape = MPM_SHAPE_BSPLINE
    else
       prob%shape = MPM_SHAPE_LINEAR
    end if
    if (USF .and. MUSL) stop 'update_step_hip: USF together with MUSL is not supported'
    if (USF) then
       prob%scheme = MPM_USF
    else if (MUSL) then
       prob%scheme = MPM_MUSL
    else
       prob%scheme = MPM_USL
    end if
    prob%quasi_load = merge(1, 0, QuasiLoad)
    prob%quasi_damp = merge(1, 0, QuasiDamp)
    prob%gravity = merge(1, 0, Gravity)
    prob%update_v_by_fp = merge(1, 0, UpdateVbyFp)
    prob%fix = FixS
    prob%ncx = NumCellx; prob%ncy = NumCelly; prob%ncz = NumCellz
    prob%dr_damping = merge(1, 0, DR_DAMPING)
    prob%smooth_stress = merge(1, 0, SmoothStress)
    prob%nb_component = nb_component
    prob%tlmpm = merge(1, 0, TLMPM)
    prob%contact = 0
    if (contact) prob%contact = Contact_type
    prob%normbody = normbody
    prob%reserved0 = 0
    prob%fricfa = fricfa
    prob%span = [SpanX(1), SpanX(2), SpanY(1), SpanY(2), SpanZ(1), SpanZ(2)]
    prob%dcell = DCell
    prob%cutoff = CutOff
    prob%dt = DT
    prob%dtscale = DTScale
    prob%endtime = EndTime
    prob%quasi_load_time = QuasiLoadTime
    prob%damp_factor = DampFactor
    prob%bq1 = bq1
    prob%bq2 = bq2
    allocate(mats(nb_mat), bodies(nb_body), ps(PF_COUNT * n), pbody(n))
    do m = 1, nb_mat
       mats(m)%mtype = mat_list(m)%MatType
       mats(m)%eos_type = mat_list(m)%EosType
       mats(m)%density = mat_list(m)%Density
       mats(m)%young = mat_list(m)%Young
       mats(m)%poisson = mat_list(m)%Poisson
       mats(m)%yield0 = mat_list(m)%Yield0
       mats(m)%tangmod = mat_list(m)%TangMod
       mats(m)%miu = mat_list(m)%miu
       mats(m)%bulk_modulus = mat_list(m)%BulkModulus
       mats(m)%wavespd = mat_list(m)%Wavespd
       mats(m)%q_fai = mat_list(m)%q_fai
       mats(m)%k_fai = mat_list(m)%k_fai
       mats(m)%q_psi = mat_list(m)%q_psi
       mats(m)%ten_f = mat_list(m)%ten_f
       mats(m)%ceos = mat_list(m)%cEos
       mats(m)%b_jc = mat_list(m)%B_jc
       mats(m)%n_jc = mat_list(m)%n_jc
       mats(m)%c_jc = mat_list(m)%C_jc
       mats(m)%m_jc = mat_list(m)%m_jc
       mats(m)%epso = mat_list(m)%epso
       mats(m)%epf = mat_list(m)%epf
       mats(m)%roomt = mat_list(m)%roomt
       mats(m)%melt = mat_list(m)%Melt
       mats(m)%spec_heat = mat_list(m)%SpecHeat
       ! SetDT/UpdateDT take max(c, D, Wavespd) (DataIn.f90:1289-1290), but the
       ! input phase sets D only for HE burn and Wavespd and miu only for the
       ! null material; otherwise the reference reads whatever the allocation
       ! held.  Those undefined values are replaced by 0 here.
       mats(m)%det_speed = 0d0
       if (mat_list(m)%MatType == 8) mats(m)%det_speed = mat_list(m)%D
       if (mat_list(m)%MatType /= 9) mats(m)%wavespd = 0d0
       if (mat_list(m)%MatType /= 9) mats(m)%miu = 0d0
       ! the null material reads no Young's modulus or Poisson ratio
       ! (DataIn.f90:486-491), yet SetDT/UpdateDT evaluate the elastic wave
       ! speed from them: 0 here, so its wave speed is Wavespd
       if (mat_list(m)%MatType == 9) then
          mats(m)%young = 0d0
          mats(m)%poisson = 0d0
       end if
    end do
    ! initial DT of SetDT (DataIn.f90:1278-1295) from the defined values
    dt0 = 1.0e6
    do m = 1, nb_mat
       cp0 = sqrt(mats(m)%young * (1 - mats(m)%poisson) / (1 + mats(m)%poisson) / &
                  (1 - 2 * mats(m)%poisson) / mats(m)%density)
       cp0 = max(cp0, mats(m)%det_speed)
       cp0 = max(cp0, mats(m)%wavespd)
       dt0 = min(dt0, DCell / cp0)
    end do
    DT = dt0 * DTScale
    prob%dt = DT
    do b = 1, nb_body
       bodies(b)%mat = body_list(b)%mat
       bodies(b)%com = body_list(b)%comID
       bodies(b)%par_begin = body_list(b)%par_begin - 1
       bodies(b)%par_end = body_list(b)%par_end
       bodies(b)%grav = body_list(b)%Gravp
       do i = body_list(b)%par_begin, body_list(b)%par_end
          pbody(i) = b - 1
       end do
    end do
    call pack_particles()
    prob%mats = c_loc(mats)
    prob%bodies = c_loc(bodies)
    prob%pstate = c_loc(ps)
    prob%pbody = c_loc(pbody)
    prec = MPM_PREC_FP64
    call get_environment_variable('MPM3D_PRECISION', env)
    if (trim(env) == 'fp32') prec = MPM_PREC_FP32
    if (trim(env) == 'fp64exact') prec = MPM_PREC_FP64_EXACT
    jc = .false.
    do m = 1, nb_mat
       jc = jc .or. mat_list(m)%MatType == 4 .or. mat_list(m)%MatType == 7
    end do
    if ((contact .or. TLMPM .or. jc) .and. prec /= MPM_PREC_FP64_EXACT) then
       write(*, '(a)') ' update_step_hip: contact, total-Lagrangian and Johnson-Cook (john, jcf) decks run in ' // &
                       'the reference-order FP64 mode'
       prec = MPM_PREC_FP64_EXACT
    end if
    call mpm_check(mpm_ctx_create(c_loc(prob), prec, ctx), 'mpm_ctx_create')
    j = 0
  end subroutine ensure_context

  subroutine pack_particles()
    use ParticleData
    integer :: i, n, a, c
    n = nb_particle
    do i = 1, n
       associate (p => particle_list(i))
         ps(PF_X * n + i) = p%Xp(1); ps(PF_Y * n + i) = p%Xp(2); ps(PF_Z * n + i) = p%Xp(3)
         ps(PF_VX * n + i) = p%VXp(1); ps(PF_VY * n + i) = p%VXp(2); ps(PF_VZ * n + i) = p%VXp(3)
         ps(PF_FX * n + i) = p%FXp(1); ps(PF_FY * n + i) = p%FXp(2); ps(PF_FZ * n + i) = p%FXp(3)
         ps(PF_MASS * n + i) = p%mass
         ps(PF_VOL * n + i) = p%VOL
         ps(PF_SM * n + i) = p%SM
         ps(PF_SEQV * n + i) = p%Seqv
         ps(PF_SDXX * n + i) = p%SDxx; ps(PF_SDYY * n + i) = p%SDyy; ps(PF_SDZZ * n + i) = p%SDzz
         ps(PF_SDYZ * n + i) = p%SDyz; ps(PF_SDXZ * n + i) = p%SDxz; ps(PF_SDXY * n + i) = p%SDxy
         ps(PF_SIGY * n + i) = p%sig_y
         ps(PF_EPEFF * n + i) = p%epeff
         ps(PF_IE * n + i) = p%ie
         ps(PF_Q * n + i) = p%q
         ps(PF_CP * n + i) = p%cp
         do a = 1, 3
            do c = 1, 3
               ps((PF_FP + 3 * (a - 1) + (c - 1)) * n + i) = p%Fp(a, c)
            end do
         end do
         ps((PF_XX + 0) * n + i) = p%XX(1); ps((PF_XX + 1) * n + i) = p%XX(2); ps((PF_XX + 2) * n + i) = p%XX(3)
         ps(PF_TEMP * n + i) = p%celsius_t
         ps(PF_DMG * n + i) = p%DMG
         ps(PF_LT * n + i) = p%LT
         ps(PF_FAIL * n + i) = merge(1d0, 0d0, p%failure)
         ps(PF_SPRE * n + i) = p%Spre
         ps((PF_XO + 0) * n + i) = p%Xo(1); ps((PF_XO + 1) * n + i) = p%Xo(2); ps((PF_XO + 2) * n + i) = p%Xo(3)
         ps(PF_VOL0 * n + i) = p%VOL0
         do a = 1, 9
            ps((PF_PK + a - 1) * n + i) = p%PK(a)
         end do
       end associate
    end do
  end subroutine pack_particles

  ! device -> particle_list (after the last particle phase of a step)
  subroutine unpack_particles()
    use ParticleData
    integer :: i, n, a, c
    n = nb_particle
    call mpm_check(mpm_download_particles(ctx, ps), 'mpm_download_particles')
    do i = 1, n
       associate (p => particle_list(i))
         p%Xp = [ps(PF_X * n + i), ps(PF_Y * n + i), ps(PF_Z * n + i)]
         p%XX = [ps(PF_XX * n + i), ps((PF_XX + 1) * n + i), ps((PF_XX + 2) * n + i)]
         p%VXp = [ps(PF_VX * n + i), ps(PF_VY * n + i), ps(PF_VZ * n + i)]
         p%VOL = ps(PF_VOL * n + i)
         p%SM = ps(PF_SM * n + i)
         p%Seqv = ps(PF_SEQV * n + i)
         p%SDxx = ps(PF_SDXX * n + i); p%SDyy = ps(PF_SDYY * n + i); p%SDzz = ps(PF_SDZZ * n + i)
         p%SDyz = ps(PF_SDYZ * n + i); p%SDxz = ps(PF_SDXZ * n + i); p%SDxy = ps(PF_SDXY * n + i)
         p%sig_y = ps(PF_SIGY * n + i)
         p%epeff = ps(PF_EPEFF * n + i)
         p%ie = ps(PF_IE * n + i)
         p%q = ps(PF_Q * n + i)
         p%cp = ps(PF_CP * n + i)
         do a = 1, 3
            do c = 1, 3
               p%Fp(a, c) = ps((PF_FP + 3 * (a - 1) + (c - 1)) * n + i)
            end do
         end do
         p%celsius_t = ps(PF_TEMP * n + i)
         p%DMG = ps(PF_DMG * n + i)
         p%failure = ps(PF_FAIL * n + i) /= 0d0
         p%Spre = ps(PF_SPRE * n + i)
         do a = 1, 9
            p%PK(a) = ps((PF_PK + a - 1) * n + i)
         end do
       end associate
    end do
  end subroutine unpack_particles

  ! what the driver reads after UpdateDT (see the module header)
  subroutine refresh_host_view()
    use ParticleData
    use GridData, only: tot_cont_for, Contact_type
    use DataOut, only: nCurves, CurvePoint, WriteTecPlot, WriteParaView, OutTime
    real(c_double) :: sc(8)
    character(len=8) :: f
    integer :: q, k, n, a, c
    ! the device's calcEnergy (after its damping, if any)
    call mpm_check(mpm_get_scalars(ctx, sc), 'mpm_get_scalars')
    EngKinetic = sc(4)
    EngInternal = sc(5)
    Momentum = sc(6:8)
    if (contact .and. Contact_type == 1) &
       call mpm_check(mpm_get_contact_force(ctx, tot_cont_for), 'mpm_get_contact_force')
    if (.not. plt_started) then
       plt_mirror = plt_mirror + OutTime
       plt_started = .true.
    end if
    call get_environment_variable('MPM3D_DUMP', f)
    full_view = .false.
    if (len_trim(f) > 0 .or. ((WriteTecPlot .or. WriteParaView) .and. CurrentTime >= plt_mirror)) then
       if ((WriteTecPlot .or. WriteParaView) .and. CurrentTime >= plt_mirror) plt_mirror = plt_mirror + OutTime
       full_view = .true.
       call unpack_particles()
       n = nb_particle
       if (.not. allocated(icell_buf)) allocate(icell_buf(n))
       call mpm_check(mpm_download_icell(ctx, icell_buf), 'mpm_download_icell')
       do q = 1, n
          particle_list(q)%icell = icell_buf(q)
       end do
       return
    end if
    k = nCurves
    if (k <= 0) return
    if (.not. allocated(sub)) allocate(sub(PF_COUNT * k))
    call mpm_check(mpm_download_particle_subset(ctx, int(k, c_int), int(CurvePoint(1:k), c_int32_t), sub), &
                   'mpm_download_particle_subset')
    do q = 1, k
       associate (p => particle_list(CurvePoint(q)))
         p%Xp = [sub(PF_X * k + q), sub(PF_Y * k + q), sub(PF_Z * k + q)]
         p%VXp = [sub(PF_VX * k + q), sub(PF_VY * k + q), sub(PF_VZ * k + q)]
         p%VOL = sub(PF_VOL * k + q)
         p%SM = sub(PF_SM * k + q)
         p%Seqv = sub(PF_SEQV * k + q)
         p%SDxx = sub(PF_SDXX * k + q); p%SDyy = sub(PF_SDYY * k + q); p%SDzz = sub(PF_SDZZ * k + q)
         p%SDyz = sub(PF_SDYZ * k + q); p%SDxz = sub(PF_SDXZ * k + q); p%SDxy = sub(PF_SDXY * k + q)
         p%sig_y = sub(PF_SIGY * k + q)
         p%epeff = sub(PF_EPEFF * k + q)
         p%ie = sub(PF_IE * k + q)
         p%q = sub(PF_Q * k + q)
         p%cp = sub(PF_CP * k + q)
         do a = 1, 3
            do c = 1, 3
               p%Fp(a, c) = sub((PF_FP + 3 * (a - 1) + (c - 1)) * k + q)
            end do
         end do
         p%celsius_t = sub(PF_TEMP * k + q)
         p%DMG = sub(PF_DMG * k + q)
         p%failure = sub(PF_FAIL * k + q) /= 0d0
       end associate
    end do
  end subroutine refresh_host_view

  subroutine maybe_dump()
    use ParticleData
    character(len=1024) :: f
    real(c_double) :: sc(8)
    integer :: u
    call get_environment_variable('MPM3D_DUMP', f)
    if (len_trim(f) == 0) return
    call mpm_check(mpm_get_scalars(ctx, sc), 'mpm_get_scalars')
    open(newunit=u, file=trim(f), access='stream', form='unformatted', status='replace')
    write(u) int(nb_particle, 4), int(istep, 4), sc
    write(u) ps
    close(u)
  end subroutine maybe_dump

end module update_step_hip_state


subroutine GridMomentumInitial()
  use update_step_hip_state
  implicit none
  call ensure_context()
  call mpm_check(mpm_begin_step(ctx), 'begin_step')   ! the driver already advanced istep/CurrentTime
  call mpm_check(mpm_grid_momentum_initial(ctx), 'GridMomentumInitial')
end subroutine GridMomentumInitial

subroutine ApplyBoundaryConditions()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_apply_boundary_conditions(ctx), 'ApplyBoundaryConditions')
end subroutine ApplyBoundaryConditions

subroutine ParticleStressUpdate()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_particle_stress_update(ctx), 'ParticleStressUpdate')
end subroutine ParticleStressUpdate

subroutine GridMomentumUpdate()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_grid_momentum_update(ctx), 'GridMomentumUpdate')
end subroutine GridMomentumUpdate

subroutine IntegrateMomentum()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_integrate_momentum(ctx), 'IntegrateMomentum')
end subroutine IntegrateMomentum

subroutine ParticlePositionUpdate()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_particle_position_update(ctx), 'ParticlePositionUpdate')
end subroutine ParticlePositionUpdate

subroutine GridMomentumMUSL()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_grid_momentum_musl(ctx), 'GridMomentumMUSL')
end subroutine GridMomentumMUSL

subroutine UpdateDT()
  use update_step_hip_state
  use ParticleData, only: DT
  implicit none
  integer(c_int), parameter :: MPM_ERR_NUMERIC = -8
  real(c_double) :: dt_dev
  integer(c_int) :: rc
  logical, save :: warned = .false.
  real(c_double) :: ek, ei, mom(3)
  ! calcEnergy (Particle.f90:204-230) ahead of UpdateDT, as the driver calls
  ! them: the driver's own calcEnergy ran on the host view; the device's
  ! energies are the ones dynamicRelaxationDamping (drda) compares
  call mpm_check(mpm_calc_energy(ctx, ek, ei, mom), 'mpm_calc_energy')
  rc = mpm_update_dt(ctx, dt_dev)
  if (rc == MPM_ERR_NUMERIC) then
     ! a non-finite volume or stress: warned once, the run goes on as the
     ! reference's would (it warns only on negative volume)
     if (.not. warned) write(*, '(a)') ' === warning: non-finite particle volume or stress on the device'
     warned = .true.
  else
     call mpm_check(rc, 'UpdateDT')
  end if
  DT = dt_dev
  call refresh_host_view()
  nsteps_done = nsteps_done + 1
  call maybe_dump()
end subroutine UpdateDT

subroutine GridMassInitial()
  use update_step_hip_state
  implicit none
  call ensure_context()
end subroutine GridMassInitial

subroutine TLGridMomentumInitial()
  use update_step_hip_state
  implicit none
  call ensure_context()
  call mpm_check(mpm_begin_step(ctx), 'begin_step')
  call mpm_check(mpm_grid_momentum_initial(ctx), 'TLGridMomentumInitial')
end subroutine TLGridMomentumInitial

subroutine TLParticleStressUpdate()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_particle_stress_update(ctx), 'TLParticleStressUpdate')
end subroutine TLParticleStressUpdate

subroutine TLGridMomentumUpdate()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_grid_momentum_update(ctx), 'TLGridMomentumUpdate')
end subroutine TLGridMomentumUpdate

subroutine TLSmoothStressbyGrid()
  implicit none
  call SmoothStressbyGrid()
end subroutine TLSmoothStressbyGrid

subroutine Lagr_NodContact()
  use update_step_hip_state
  implicit none
  call mpm_check(mpm_lagr_nod_contact(ctx), 'Lagr_NodContact')
end subroutine Lagr_NodContact

subroutine SmoothStressbyGrid()
  use update_step_hip_state
  use ParticleData, only: particle_list, nb_particle
  implicit none
  integer :: i, n
  call mpm_check(mpm_smooth_stress_by_grid(ctx), 'SmoothStressbyGrid')
  if (.not. full_view) return
  ! an output step: the writers read Spre of every particle
  n = nb_particle
  call mpm_check(mpm_download_particles(ctx, ps), 'mpm_download_particles')
  do i = 1, n
     particle_list(i)%Spre = ps(PF_SPRE * n + i)
  end do
  call maybe_dump()
end subroutine SmoothStressbyGrid
