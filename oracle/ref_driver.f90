! ---------------------------------------------------------------------------
! ref_driver -- dump driver linked against the reference's own modules and
! step routines (compiled from /root/reference/src by oracle/build_ref.sh).
!
! Test infrastructure only.  It replays the updated-Lagrangian step sequence
! of the reference main program (MPM3D.f90:69-113: GridMomentumInitial,
! [USF: ApplyBoundaryConditions, ParticleStressUpdate], GridMomentumUpdate,
! IntegrateMomentum, [cont lagr: Lagr_NodContact], ParticlePositionUpdate,
! [MUSL: GridMomentumMUSL,
! ApplyBoundaryConditions], [USL/MUSL: ParticleStressUpdate], calcEnergy,
! UpdateDT, [smoo: SmoothStressbyGrid], [drda: dynamicRelaxationDamping])
! without the per-step file output, and writes the complete particle and
! grid state as little-endian binary after chosen steps.  A total-Lagrangian
! deck (TLmp) replays MPM3D.f90:164-250 instead: GridMassInitial once, then
! TLGridMomentumInitial, [USF: ApplyBoundaryConditions,
! TLParticleStressUpdate], TLGridMomentumUpdate, IntegrateMomentum,
! ParticlePositionUpdate, [MUSL ...], [USL/MUSL: TLParticleStressUpdate],
! calcEnergy, UpdateDT, [smoo: TLSmoothStressbyGrid], [drda].
!
! usage:  ref_driver <deck> <nsteps> <outprefix> <dumpsteps> [phasestep]
!   <deck>       .mpm input deck (read by the reference's InputPara)
!   <nsteps>     number of steps to run (the deck's endtime is ignored)
!   <dumpsteps>  comma list of steps after which <outprefix>_sNNNNN.bin is
!                written; step 0 is the state right after InputPara+calcEnergy
!   [phasestep]  if given, every phase of that step is also dumped to
!                <outprefix>_pNNNNN_K.bin (K = phase tag, see below)
!
! Wall time of the step loop alone is printed as
!   "STEPLOOP <seconds> <particles> <steps>"
! so the same binary doubles as the timed reference CPU baseline.
! ---------------------------------------------------------------------------
program ref_driver
  use ParticleData
  use GridData
  use MaterialData
  use DataIn
  implicit none

  character(len=512) :: deck, prefix, arg
  integer :: nsteps, phasestep, k, n, pos
  integer :: dumpat(256), ndump
  integer(8) :: c0, c1, crate
  logical :: dophase

  external GridMomentumInitial, ParticleStressUpdate, GridMomentumUpdate, &
           IntegrateMomentum, ParticlePositionUpdate, GridMomentumMUSL, &
           ApplyBoundaryConditions, UpdateDT, SmoothStressbyGrid, Lagr_NodContact, &
           GridMassInitial, TLGridMomentumInitial, TLParticleStressUpdate, TLGridMomentumUpdate, &
           TLSmoothStressbyGrid

  if (command_argument_count() < 4) then
     print *, 'usage: ref_driver <deck> <nsteps> <outprefix> <dumpsteps> [phasestep]'
     stop 2
  end if
  call get_command_argument(1, deck)
  call get_command_argument(2, arg);  read(arg, *) nsteps
  call get_command_argument(3, prefix)
  call get_command_argument(4, arg)
  ndump = 0
  do while (len_trim(arg) > 0)
     pos = index(arg, ',')
     ndump = ndump + 1
     if (pos > 0) then
        read(arg(1:pos-1), *) dumpat(ndump)
        arg = arg(pos+1:)
     else
        read(arg, *) dumpat(ndump)
        arg = ' '
     end if
  end do
  phasestep = -1
  if (command_argument_count() >= 5) then
     call get_command_argument(5, arg);  read(arg, *) phasestep
  end if

  ! InputPara reads the deck named by the first command-line argument.
  call InputPara()
  call calcEnergy()

  if (any(dumpat(1:ndump) == 0)) call dump_state(0, 0)

  if (TLMPM) call GridMassInitial()
  call system_clock(c0, crate)
  do k = 1, nsteps
     istep = istep + 1
     CurrentTime = CurrentTime + DT
     EngInternal = 0.0
     dophase = (istep == phasestep)

     if (TLMPM) then
        call TLGridMomentumInitial()
     else
        call GridMomentumInitial()
     end if
     if (dophase) call dump_state(istep, 1)
     if (USF) then
        call ApplyBoundaryConditions()
        if (TLMPM) then
           call TLParticleStressUpdate()
        else
           call ParticleStressUpdate()
        end if
        if (dophase) call dump_state(istep, 2)
     end if
     if (TLMPM) then
        call TLGridMomentumUpdate()
     else
        call GridMomentumUpdate()
     end if
     if (dophase) call dump_state(istep, 3)
     call IntegrateMomentum()
     if (Contact_type == 1) call Lagr_NodContact()
     if (dophase) call dump_state(istep, 4)
     call ParticlePositionUpdate()
     if (dophase) call dump_state(istep, 5)
     if (MUSL) then
        call GridMomentumMUSL()
        call ApplyBoundaryConditions()
        if (dophase) call dump_state(istep, 6)
     end if
     if (.not. USF) then
        if (TLMPM) then
           call TLParticleStressUpdate()
        else
           call ParticleStressUpdate()
        end if
        if (dophase) call dump_state(istep, 7)
     end if
     call calcEnergy()
     call UpdateDT()
     if (SmoothStress) then
        if (TLMPM) then
           call TLSmoothStressbyGrid()
        else
           call SmoothStressbyGrid()
        end if
     end if
     if (DR_DAMPING) call dynamicRelaxationDamping()
     if (dophase) call dump_state(istep, 8)

     do n = 1, ndump
        if (dumpat(n) == istep) call dump_state(istep, 0)
     end do
  end do
  call system_clock(c1)
  write(*, '(a,es16.8,2i12)') 'STEPLOOP ', dble(c1 - c0) / dble(crate), nb_particle, nsteps

contains

  subroutine dump_state(step, phase)
    integer, intent(in) :: step, phase
    character(len=600) :: fname
    integer :: u, i, j, c, ncomp
    integer(4) :: hdr(16)
    real(8) :: rh(16)

    if (phase == 0) then
       write(fname, '(a,"_s",i5.5,".bin")') trim(prefix), step
    else
       write(fname, '(a,"_p",i5.5,"_",i1,".bin")') trim(prefix), step, phase
    end if
    open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')

    hdr = 0
    hdr(1) = 1146113101          ! 'MPMD' tag
    hdr(2) = 5                   ! layout version
    hdr(3) = nb_particle
    hdr(4) = nb_gridnode
    hdr(5) = istep
    hdr(6) = phase
    hdr(7) = NumCellx;  hdr(8) = NumCelly;  hdr(9) = NumCellz
    hdr(10) = nb_body;  hdr(11) = nb_mat
    ncomp = 1
    if (contact) ncomp = nb_component
    hdr(12) = ncomp
    rh = 0d0
    rh(1) = DT;  rh(2) = CurrentTime;  rh(3) = EngKinetic;  rh(4) = EngInternal
    rh(5:7) = Momentum;  rh(8) = CutOff;  rh(9) = DCell
    rh(10) = SpanX(1); rh(11) = SpanX(2); rh(12) = SpanY(1); rh(13) = SpanY(2)
    rh(14) = SpanZ(1); rh(15) = SpanZ(2); rh(16) = DTScale
    write(u) hdr, rh

    ! particle arrays, one field after another (oracle/pyoracle.py: read_ref_dump)
    do j = 1, 3
       write(u) (particle_list(i)%Xp(j), i = 1, nb_particle)
    end do
    do j = 1, 3
       write(u) (particle_list(i)%XX(j), i = 1, nb_particle)
    end do
    do j = 1, 3
       write(u) (particle_list(i)%VXp(j), i = 1, nb_particle)
    end do
    do j = 1, 3
       write(u) (particle_list(i)%FXp(j), i = 1, nb_particle)
    end do
    write(u) (particle_list(i)%mass, i = 1, nb_particle)
    write(u) (particle_list(i)%VOL, i = 1, nb_particle)
    write(u) (particle_list(i)%SM, i = 1, nb_particle)
    write(u) (particle_list(i)%Seqv, i = 1, nb_particle)
    write(u) (particle_list(i)%SDxx, i = 1, nb_particle)
    write(u) (particle_list(i)%SDyy, i = 1, nb_particle)
    write(u) (particle_list(i)%SDzz, i = 1, nb_particle)
    write(u) (particle_list(i)%SDyz, i = 1, nb_particle)
    write(u) (particle_list(i)%SDxz, i = 1, nb_particle)
    write(u) (particle_list(i)%SDxy, i = 1, nb_particle)
    write(u) (particle_list(i)%sig_y, i = 1, nb_particle)
    write(u) (particle_list(i)%epeff, i = 1, nb_particle)
    write(u) (particle_list(i)%ie, i = 1, nb_particle)
    write(u) (particle_list(i)%q, i = 1, nb_particle)
    write(u) (particle_list(i)%cp, i = 1, nb_particle)
    write(u) (particle_list(i)%celsius_t, i = 1, nb_particle)
    write(u) (particle_list(i)%DMG, i = 1, nb_particle)
    write(u) (particle_list(i)%LT, i = 1, nb_particle)
    write(u) (merge(1d0, 0d0, particle_list(i)%failure), i = 1, nb_particle)
    write(u) (particle_list(i)%Spre, i = 1, nb_particle)
    do j = 1, 3
       write(u) (particle_list(i)%Xo(j), i = 1, nb_particle)
    end do
    write(u) (particle_list(i)%VOL0, i = 1, nb_particle)
    do j = 1, 9
       write(u) (particle_list(i)%PK(j), i = 1, nb_particle)
    end do
    do j = 1, 3          ! Fp(i,j) column-major: (1,1),(2,1),(3,1),(1,2),...
       do n = 1, 3
          write(u) (particle_list(i)%Fp(n, j), i = 1, nb_particle)
       end do
    end do
    write(u) (dble(particle_list(i)%icell), i = 1, nb_particle)
    ! grid arrays of component 1
    write(u) (grid_list(1, i)%Mg, i = 1, nb_gridnode)
    do j = 1, 3
       write(u) (grid_list(1, i)%PXg(j), i = 1, nb_gridnode)
    end do
    do j = 1, 3
       write(u) (grid_list(1, i)%FXg(j), i = 1, nb_gridnode)
    end do
    ! node fixities (1.0 = fixed) for x, y, z
    write(u) (merge(1d0, 0d0, node_list(i)%Fix_x), i = 1, nb_gridnode)
    write(u) (merge(1d0, 0d0, node_list(i)%Fix_y), i = 1, nb_gridnode)
    write(u) (merge(1d0, 0d0, node_list(i)%Fix_z), i = 1, nb_gridnode)
    ! grids of the other components (contact), then tot_cont_for
    do c = 2, ncomp
       write(u) (grid_list(c, i)%Mg, i = 1, nb_gridnode)
       do j = 1, 3
          write(u) (grid_list(c, i)%PXg(j), i = 1, nb_gridnode)
       end do
       do j = 1, 3
          write(u) (grid_list(c, i)%FXg(j), i = 1, nb_gridnode)
       end do
    end do
    if (Contact_type == 1 .and. istep > 0) then
       write(u) tot_cont_for
    else
       write(u) 0d0, 0d0, 0d0
    end if
    close(u)
  end subroutine dump_state

end program ref_driver
