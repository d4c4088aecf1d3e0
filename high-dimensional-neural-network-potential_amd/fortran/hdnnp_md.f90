! Fortran driver of the MI355X HDNNP engine, written against the reference's
! public interface (modules hdnnps / atomicStructure / constants).
!
!   hdnnp_md md   <paramdir/> <xyz> <momenta> <nsteps> <out> [mixed]
!       velocity Verlet of hdnnpExample.f90:58-72 (dt = 0.5 fs), initial
!       momenta read from a file (the reference draws unseeded random ones)
!   hdnnp_md eval <paramdir/> <structure> <out> [mixed]
!       one hdnnpEnergyAndForces call with dEdLat, plus dEdLat2Stress
!   hdnnp_md committee <nnps.txt> <structure> <out>
!       one globalHdnnpEnergyAndForces call (committee mean, energy std)
!   hdnnp_md stages <paramdir/> <structure> <T|F> <ndsf> <out> [<paramdir2/>|-] [mixed]
!       the stage routines in the order of globalHdnnpEnergyAndForces
!       (globalHdnnps.f90:47-55): hdnnpCalcNeighborLists and
!       hdnnpCalcSymfunctions (calcSfsTogether T or F) of the first
!       potential, then hdnnpCalcNNs of it and, given a second directory, of
!       that potential on the first one's symmetry functions (lines energy2,
!       force2, dEdLat2). Written as refdump stages writes the reference's.
!   hdnnp_md repeat <structure> <n1> <n2> <n3> <out>
!       as_repeatUnitCell of a structure, written as `read` writes one
!   hdnnp_md read <format> <file> <out>
!       the structure(s) of a file read by as_readXYZ / as_readFhiaims /
!       as_readAscii ('ascii', or 'asciiang' with isAngstrom) /
!       as_readPOSCAR / as_readData ('xyz', 'fhiaims', 'poscar', 'data'),
!       written as 'structure nat periodic', lattice rows, 'atom Z x y z'
!       and for data 'energy E Q' and 'aux q fx fy fz' per atom
!
! Structure file: "nat periodic" / three lattice-vector lines / "Z x y z" rows.
program hdnnp_md
    use precision
    use constants
    use atomicStructure
    use hdnnps
    use globalHdnnps
    implicit none
    character(len=512) :: mode, pdir, a3, a4, a5, a6, a7, a8
    type(hdnnp) :: h
    integer :: prec

    call get_command_argument(1, mode)
    call get_command_argument(2, pdir)
    call get_command_argument(3, a3)
    call get_command_argument(4, a4)
    call get_command_argument(5, a5)
    call get_command_argument(6, a6)
    call get_command_argument(7, a7)
    call get_command_argument(8, a8)
    prec = HDNNP_FP64
    if (trim(mode) == 'read') then
        call runRead(trim(pdir), trim(a3), trim(a4))
        stop
    end if
    if (trim(mode) == 'repeat') then
        call runRepeat(trim(pdir), trim(a3), trim(a4), trim(a5), trim(a6))
        stop
    end if
    select case (trim(mode))
    case ('md')
        if (trim(a7) == 'mixed') prec = HDNNP_MIXED
        call parseRuNNerInput(h, trim(pdir), prec)
        call runMD(h, trim(a3), trim(a4), trim(a5), trim(a6))
    case ('eval')
        if (trim(a5) == 'mixed') prec = HDNNP_MIXED
        call parseRuNNerInput(h, trim(pdir), prec)
        call runEval(h, trim(a3), trim(a4))
    case ('committee')
        call initGlobalHdnnpsFromInputFile(file=trim(pdir))
        call runCommittee(trim(a3), trim(a4))
    case ('stages')
        if (trim(a8) == 'mixed') prec = HDNNP_MIXED
        call initHdnnp(h)
        call parseRuNNerInput(h, trim(pdir), prec)
        call runStages(h, trim(a3), trim(a4) == 'T', trim(a5), trim(a6), trim(a7), prec)
    case default
        print*, 'usage: hdnnp_md md|eval|committee ...'
        error stop 2
    end select
    call hdnnpDestroy(h)

contains

    subroutine writeStruct(u, ats, aux)
        integer, intent(in) :: u
        type(atStruct), intent(in) :: ats
        logical, intent(in) :: aux
        integer :: i
        write(u, '(A,I8,I2)') 'structure ', ats%nat, merge(1, 0, ats%periodic)
        do i = 1, 3
            write(u, '(3ES25.16E3)') merge(ats%lat(:, i), 0._dp, ats%periodic)
        end do
        do i = 1, ats%nat
            write(u, '(A,I4,3ES25.16E3)') 'atom ', ats%el(i), ats%ats(:, i)
        end do
        if (aux) then
            write(u, '(A,2ES25.16E3)') 'energy ', ats%energy, ats%qtot
            do i = 1, ats%nat
                write(u, '(A,4ES25.16E3)') 'aux ', ats%q(i), ats%f(:, i)
            end do
        end if
    end subroutine writeStruct

    subroutine runRepeat(sfile, a1, a2, a3, oname)
        character(len=*), intent(in) :: sfile, a1, a2, a3, oname
        type(atStruct) :: ats
        integer :: n(3), u
        read(a1, *) n(1)
        read(a2, *) n(2)
        read(a3, *) n(3)
        call readStructure(sfile, ats)
        call as_repeatUnitCell(ats, n)
        open(newunit=u, file=oname, action='write', status='replace')
        call writeStruct(u, ats, .false.)
        close(u)
    end subroutine runRepeat

    subroutine runRead(fmt, fname, oname)
        character(len=*), intent(in) :: fmt, fname, oname
        type(atStruct) :: ats
        type(atStruct), allocatable :: many(:)
        integer :: u, n, i
        open(newunit=u, file=oname, action='write', status='replace')
        select case (fmt)
        case ('xyz')
            call as_readXYZ(fname, ats)
        case ('fhiaims')
            call as_readFhiaims(fname, ats)
        case ('ascii')
            call as_readAscii(fname, ats)
        case ('asciiang')
            call as_readAscii(fname, ats, .true.)
        case ('poscar')
            call as_readPOSCAR(fname, ats)
        case ('data')
            call as_readData(fname, n, many)
            do i = 1, n
                call writeStruct(u, many(i), .true.)
            end do
            close(u)
            return
        case default
            error stop 'unknown format'
        end select
        call writeStruct(u, ats, .false.)
        close(u)
    end subroutine runRead

    subroutine runMD(h, xyz, pfile, nstr, fname)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: xyz, pfile, nstr, fname
        type(atStruct) :: ats
        real(dp), allocatable :: forces(:,:), m(:,:), p(:,:), lastForces(:,:)
        real(dp), parameter :: stepsize = 0.5_dp * femtoseconds2atomic
        real(dp) :: epot, ekin
        integer :: i, u, nsteps

        read(nstr, *) nsteps
        call as_readXYZ(xyz, ats)
        allocate(forces(3, ats%nat), lastForces(3, ats%nat), m(3, ats%nat), p(3, ats%nat))
        m(1,:) = getAtomicMass(ats%el) * massUnit2ElectronMass
        m(2,:) = m(1,:)
        m(3,:) = m(1,:)
        open(newunit=u, file=pfile, action='read', status='old')
        do i = 1, ats%nat
            read(u, *) p(:, i)
        end do
        close(u)

        open(newunit=u, file=fname, action='write', status='replace')
        call hdnnpEnergyAndForces(h, ats, epot, forces)
        write(u, '(A,I6,ES25.16E3)') 'init ', -1, epot
        do i = 0, nsteps - 1
            ats%ats = ats%ats + p / m * stepSize + 0.5_dp * forces / m * stepSize**2
            lastForces = forces
            call hdnnpEnergyAndForces(h, ats, epot, forces)
            p = p + stepSize / 2._dp * (forces + lastForces)
            ekin = sum(p**2 / m) / 2._dp
            write(u, '(A,I6,3ES25.16E3)') 'step ', i, epot, ekin, epot + ekin
        end do
        write(u, '(A)') 'positions'
        do i = 1, ats%nat
            write(u, '(3ES25.16E3)') ats%ats(:, i)
        end do
        close(u)
    end subroutine runMD

    subroutine readStructure(sfile, ats)
        character(len=*), intent(in) :: sfile
        type(atStruct), intent(out) :: ats
        integer :: u, i, iper
        open(newunit=u, file=sfile, action='read', status='old')
        read(u, *) ats%nat, iper
        ats%periodic = (iper /= 0)
        do i = 1, 3
            read(u, *) ats%lat(:, i)
        end do
        allocate(ats%ats(3, ats%nat), ats%el(ats%nat))
        do i = 1, ats%nat
            read(u, *) ats%el(i), ats%ats(:, i)
        end do
        close(u)
    end subroutine readStructure

    subroutine runCommittee(sfile, fname)
        character(len=*), intent(in) :: sfile, fname
        type(atStruct) :: ats
        real(dp) :: energy, dEdLat(3,3), sd
        real(dp), allocatable :: force(:,:)
        integer :: u, i
        call readStructure(sfile, ats)
        allocate(force(3, ats%nat))
        call globalHdnnpEnergyAndForces(ats, energy, force, dEdLat, sd)
        open(newunit=u, file=fname, action='write', status='replace')
        write(u, '(A,I4)') 'members ', nGlobalHdnnps
        write(u, '(A,ES25.16E3)') 'energy ', energy
        write(u, '(A,ES25.16E3)') 'std ', sd
        write(u, '(A)') 'force'
        do i = 1, ats%nat
            write(u, '(3ES25.16E3)') force(:, i)
        end do
        write(u, '(A)') 'dEdLat'
        do i = 1, 3
            write(u, '(3ES25.16E3)') dEdLat(i, :)
        end do
        close(u)
    end subroutine runCommittee

    subroutine runStages(h, sfile, together, nd, fname, pdir2, prec)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: sfile, nd, fname, pdir2
        logical, intent(in) :: together
        integer, intent(in) :: prec
        type(atStruct) :: ats
        type(hdnnp) :: h2
        type(neighborList), allocatable :: nl(:)
        type(symFunctionContainer), allocatable :: sfs(:)
        real(dp) :: energy, dEdLat(3,3)
        real(dp), allocatable :: force(:,:)
        integer :: u, i, n, s, ndsf

        read(nd, *) ndsf
        call readStructure(sfile, ats)
        allocate(force(3, ats%nat), nl(ats%nat), sfs(ats%nat))
        call hdnnpCalcNeighborLists(h, ats, nl)
        call hdnnpCalcSymfunctions(h, ats, nl, sfs, together)
        call hdnnpCalcNNs(h, ats, nl, sfs, energy, force, dEdLat)
        open(newunit=u, file=fname, action='write', status='replace')
        write(u, '(A,I8)') 'stages ', ats%nat
        write(u, '(A,ES25.16E3)') 'energy ', energy
        write(u, '(A)') 'force'
        do i = 1, ats%nat
            write(u, '(3ES25.16E3)') force(:, i)
        end do
        write(u, '(A)') 'dEdLat'
        do i = 1, 3
            write(u, '(3ES25.16E3)') dEdLat(i, :)
        end do
        do i = 1, ats%nat
            write(u, '(A,3I8)') 'atom ', i, nl(i)%nNeis, sfs(i)%nsf
            write(u, '(*(I8))') nl(i)%ineis(1:nl(i)%nNeis)
            do n = 1, nl(i)%nNeis
                write(u, '(3ES25.16E3)') nl(i)%neis(:, n)
            end do
            write(u, '(*(ES25.16E3))') (sfs(i)%sfs(s)%sf, s = 1, sfs(i)%nsf)
            if (i <= ndsf) then
                do s = 1, sfs(i)%nsf
                    write(u, '(*(ES25.16E3))') sfs(i)%sfs(s)%dsf(:, 1:nl(i)%nNeis)
                end do
            end if
        end do
        if (len_trim(pdir2) > 0 .and. pdir2 /= '-') then
            ! globalHdnnps.f90:53-55: another potential's networks on these
            ! symmetry functions
            call initHdnnp(h2)
            call parseRuNNerInput(h2, pdir2, prec)
            call hdnnpCalcNNs(h2, ats, nl, sfs, energy, force, dEdLat)
            write(u, '(A,ES25.16E3)') 'energy2 ', energy
            write(u, '(A)') 'force2'
            do i = 1, ats%nat
                write(u, '(3ES25.16E3)') force(:, i)
            end do
            write(u, '(A)') 'dEdLat2'
            do i = 1, 3
                write(u, '(3ES25.16E3)') dEdLat(i, :)
            end do
            call hdnnpDestroy(h2)
        end if
        close(u)
    end subroutine runStages

    subroutine runEval(h, sfile, fname)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: sfile, fname
        type(atStruct) :: ats
        real(dp) :: energy, dEdLat(3,3), stress(3,3)
        real(dp), allocatable :: force(:,:)
        integer :: u, i, iper

        open(newunit=u, file=sfile, action='read', status='old')
        read(u, *) ats%nat, iper
        ats%periodic = (iper /= 0)
        do i = 1, 3
            read(u, *) ats%lat(:, i)
        end do
        allocate(ats%ats(3, ats%nat), ats%el(ats%nat), force(3, ats%nat))
        do i = 1, ats%nat
            read(u, *) ats%el(i), ats%ats(:, i)
        end do
        close(u)

        call hdnnpEnergyAndForces(h, ats, energy, force, dEdLat)
        stress = 0._dp
        if (ats%periodic) call dEdLat2Stress(ats%lat, dEdLat, stress)
        open(newunit=u, file=fname, action='write', status='replace')
        write(u, '(A,ES25.16E3)') 'energy ', energy
        write(u, '(A)') 'force'
        do i = 1, ats%nat
            write(u, '(3ES25.16E3)') force(:, i)
        end do
        write(u, '(A)') 'dEdLat'
        do i = 1, 3
            write(u, '(3ES25.16E3)') dEdLat(i, :)
        end do
        write(u, '(A)') 'stress'
        do i = 1, 3
            write(u, '(3ES25.16E3)') stress(i, :)
        end do
        close(u)
    end subroutine runEval

end program hdnnp_md
