! refdump -- test-infrastructure driver for the reference HDNNP (Fortran).
!
! This program is written for this repository. It is linked against the
! reference's own modules (hdnnps, neighborLists, symmetryFunctions,
! neuralNetworks, atomicStructure, ...), compiled from their source files
! where they lie by oracle/Makefile into oracle/_ref/. It is the checker's
! source of golden vectors and the "reference" CPU baseline; no product code
! calls it.
!
! Usage:
!   refdump table <paramdir/> <out>
!       parsed potential: Z-sorted elements, per element the sorted SF table
!       (hdnnps.f90:235-267), scaling (hdnnps.f90:269-300) and weights
!       (hdnnps.f90:302-330).
!   refdump eval  <paramdir/> <structure> <out>
!       hdnnpEnergyAndForces (hdnnps.f90:333-371) plus the stage outputs:
!       neighbour entries (i, j, image shift) from hdnnpCalcNeighborLists
!       (hdnnps.f90:373-388), scaled G from hdnnpCalcSymfunctions
!       (hdnnps.f90:390-485) and dE/dG from nn_forward/nn_backward
!       (neuralNetworks.f90:220-313), exactly as hdnnpCalcNNs calls them.
!   refdump md    <paramdir/> <xyzfile> <momenta> <nsteps> <out>
!       velocity-Verlet loop of hdnnpExample.f90:58-72 with given momenta.
!   refdump time  <paramdir/> <structure> <nrep> <out>
!       wall time of hdnnpEnergyAndForces (E+F+dE/dh), warm-up 1, nrep timed.
!   refdump stages <paramdir/> <structure> <T|F> <ndsf> <out>
!       the stage routines in the order hdnnpEnergyAndForces calls them:
!       hdnnpCalcNeighborLists, hdnnpCalcSymfunctions (calcSfsTogether T or
!       F) and hdnnpCalcNNs (hdnnps.f90:373-562): energy, forces, dEdLat,
!       then per atom its list (ineis, neis), its scaled G and, for the first
!       ndsf atoms, dsf of every function.
!   refdump repeat <structure> <n1> <n2> <n3> <out>
!       as_repeatUnitCell (atomicStructure.f90:636-671) of a structure.
!   refdump read  <format> <file> <out>
!       the structure(s) one of the reference's readers makes of a file:
!       format xyz (as_readXYZ), fhiaims (as_readFhiaims), ascii / asciiang
!       (as_readAscii without / with isAngstrom), poscar (as_readPOSCAR),
!       data (as_readData, atomicStructure.f90:143-475).
!
! Structure file (text): "nat periodic" / three lines, line k = lattice
! vector k / nat lines "Z x y z" (bohr).
program refdump
    use precision
    use constants
    use atomicStructure
    use pointsets
    use linalg
    use neighborLists
    use symmetryFunctions
    use neuralNetworks
    use hdnnps
    implicit none

    character(len=512) :: mode, pdir, sfile, ofile, arg
    type(hdnnp) :: h

    call get_command_argument(1, mode)
    call get_command_argument(2, pdir)
    if (trim(mode) == 'read') then
        call get_command_argument(3, sfile)
        call get_command_argument(4, ofile)
        call dumpRead(trim(pdir), trim(sfile), trim(ofile))
        stop
    end if
    if (trim(mode) == 'repeat') then
        block
            type(atStruct) :: ats
            integer :: n(3), k, u
            character(len=32) :: a
            do k = 1, 3
                call get_command_argument(2 + k, a)
                read(a, *) n(k)
            end do
            call get_command_argument(6, ofile)
            call readStruct(trim(pdir), ats)
            call as_repeatUnitCell(ats, n)
            open(newunit=u, file=trim(ofile), action='write', status='replace')
            call writeStruct(u, ats, .false.)
            close(u)
        end block
        stop
    end if
    call parseRuNNerInput(h, trim(pdir))

    select case (trim(mode))
    case ('table')
        call get_command_argument(3, ofile)
        call dumpTable(h, trim(ofile))
    case ('eval')
        call get_command_argument(3, sfile)
        call get_command_argument(4, ofile)
        call dumpEval(h, trim(sfile), trim(ofile))
    case ('md')
        call get_command_argument(3, sfile)
        call get_command_argument(4, arg)
        call get_command_argument(6, ofile)
        call runMD(h, trim(sfile), trim(arg), trim(ofile))
    case ('time')
        call get_command_argument(3, sfile)
        call get_command_argument(5, ofile)
        call runTime(h, trim(sfile), trim(ofile))
    case ('stages')
        call get_command_argument(3, sfile)
        call get_command_argument(4, arg)
        call get_command_argument(6, ofile)
        block
            character(len=32) :: nd
            integer :: ndsf
            call get_command_argument(5, nd)
            read(nd, *) ndsf
            call dumpStages(h, trim(sfile), trim(arg) == 'T', ndsf, trim(ofile))
        end block
    case default
        print*, 'unknown mode'
        stop 2
    end select

contains

    subroutine readStruct(fname, ats)
        character(len=*), intent(in) :: fname
        type(atStruct), intent(out) :: ats
        integer :: u, i, iper

        open(newunit=u, file=fname, action='read', status='old')
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
        if (.not. ats%periodic) ats%lat = 0._dp
    end subroutine readStruct

    ! one structure: 'structure nat periodic', the lattice (line k = vector
    ! k), 'atom Z x y z' lines; with aux, then 'energy E Q' and per atom
    ! 'aux q fx fy fz'
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

    subroutine dumpRead(fmt, fname, oname)
        character(len=*), intent(in) :: fmt, fname, oname
        type(atStruct) :: ats
        type(atStruct), allocatable :: many(:)
        integer :: u, n, i

        open(newunit=u, file=oname, action='write', status='replace')
        select case (fmt)
        case ('xyz')
            call as_readXYZ(fname, ats)
            call writeStruct(u, ats, .false.)
        case ('fhiaims')
            call as_readFhiaims(fname, ats)
            call writeStruct(u, ats, .false.)
        case ('ascii')
            call as_readAscii(fname, ats)
            call writeStruct(u, ats, .false.)
        case ('asciiang')
            call as_readAscii(fname, ats, .true.)
            call writeStruct(u, ats, .false.)
        case ('poscar')
            call as_readPOSCAR(fname, ats)
            call writeStruct(u, ats, .false.)
        case ('data')
            call as_readData(fname, n, many)
            do i = 1, n
                call writeStruct(u, many(i), .true.)
            end do
        case default
            print*, 'unknown format'
            stop 2
        end select
        close(u)
    end subroutine dumpRead

    subroutine dumpTable(h, fname)
        type(hdnnp), intent(in) :: h
        character(len=*), intent(in) :: fname
        integer :: u, z, s, l, i, j

        open(newunit=u, file=fname, action='write', status='replace')
        write(u, '(A,ES25.16E3)') 'maxrc ', h%maxRc
        write(u, '(A,2L2)') 'flags ', h%scalesf, h%centersf
        do z = 1, maxElemNum
            if (.not. allocated(h%sfLists(z)%sfs)) cycle
            write(u, '(A,I4,I5,ES25.16E3)') 'element ', z, h%sfLists(z)%nsf, h%atomenergies(z)
            do s = 1, h%sfLists(z)%nsf
                associate (sf => h%sfLists(z)%sfs(s))
                    write(u, '(A,I2,2I4,*(ES25.16E3))') 'sf ', sf%name, sf%sfEls(1), &
                        merge(sf%sfEls(2), 0, sf%name == 3), sf%rcut, sf%eta, &
                        merge(sf%rs, 0._dp, sf%name == 2), merge(sf%lambda, 0._dp, sf%name == 3), &
                        merge(sf%zeta, 0._dp, sf%name == 3), sf%sfmin, sf%sfmax, sf%sfmean
                end associate
            end do
            write(u, '(A,I3)') 'nlayers ', h%nnList(z)%nLayers
            do l = 1, h%nnList(z)%nLayers
                associate (ly => h%nnList(z)%layers(l))
                    write(u, '(A,2I5,1X,A1)') 'layer ', ly%nInp, ly%nOut, ly%activation
                    do j = 1, ly%nInp
                        do i = 1, ly%nOut
                            write(u, '(ES25.16E3)') ly%weights(i, j)
                        end do
                    end do
                    do i = 1, ly%nOut
                        write(u, '(ES25.16E3)') ly%bias(i)
                    end do
                end associate
            end do
        end do
        write(u, '(A,2ES25.16E3)') 'energy_range ', h%emin, h%emax
        close(u)
    end subroutine dumpTable

    subroutine dumpEval(h, sfile, fname)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: sfile, fname
        type(atStruct) :: ats, w2
        type(neighborList), allocatable :: nl(:)
        type(symFunctionContainer), allocatable :: sfs(:)
        type(neuralNetwork) :: tmpNN
        real(dp) :: energy, dEdLat(3,3), invlat(3,3), shift(3), out(1), ones(1)
        real(dp), allocatable :: force(:,:), g(:), dg(:)
        integer :: u, i, n, s, nsf, nent

        call readStruct(sfile, ats)
        allocate(force(3, ats%nat), nl(ats%nat), sfs(ats%nat))

        call hdnnpEnergyAndForces(h, ats, energy, force, dEdLat)
        if (.not. ats%periodic) dEdLat = 0._dp

        call hdnnpCalcNeighborLists(h, ats, nl)
        call hdnnpCalcSymfunctions(h, ats, nl, sfs, .true.)

        ! positions as wrapped twice by the reference (hdnnps.f90:381 and
        ! neighborLists.f90:156) to recover the integer image of each entry
        w2 = ats
        if (w2%periodic) then
            call as_moveAtomsIntoCell(w2)
            call as_moveAtomsIntoCell(w2)
            call inv3DM(w2%lat, invlat)
        end if

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
        nent = 0
        do i = 1, ats%nat
            nent = nent + nl(i)%nNeis - 1
        end do
        write(u, '(A,I12)') 'neighbors ', nent
        do i = 1, ats%nat
            do n = 2, nl(i)%nNeis
                shift = 0._dp
                if (ats%periodic) shift = matMulVec3(invlat, nl(i)%neis(:, n) - w2%ats(:, nl(i)%ineis(n)))
                write(u, '(2I8,3I4)') i - 1, nl(i)%ineis(n) - 1, nint(shift)
            end do
        end do
        write(u, '(A)') 'sf'
        do i = 1, ats%nat
            nsf = sfs(i)%nsf
            allocate(g(nsf), dg(nsf))
            do s = 1, nsf
                g(s) = sfs(i)%sfs(s)%sf
            end do
            tmpNN = h%nnList(ats%el(i))
            call nn_forward(tmpNN, g, out)
            ones(1) = 1._dp
            call nn_backward(tmpNN, ones, dg, .false.)
            write(u, '(A,2I8,ES25.16E3)') 'atom ', i - 1, nsf, out(1) + h%atomenergies(ats%el(i))
            write(u, '(*(ES25.16E3))') g
            write(u, '(*(ES25.16E3))') dg
            deallocate(g, dg)
        end do
        close(u)
    end subroutine dumpEval

    subroutine dumpStages(h, sfile, together, ndsf, fname)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: sfile, fname
        logical, intent(in) :: together
        integer, intent(in) :: ndsf
        type(atStruct) :: ats
        type(neighborList), allocatable :: nl(:)
        type(symFunctionContainer), allocatable :: sfs(:)
        real(dp) :: energy, dEdLat(3,3)
        real(dp), allocatable :: force(:,:)
        integer :: u, i, n, s

        call readStruct(sfile, ats)
        allocate(force(3, ats%nat), nl(ats%nat), sfs(ats%nat))
        call hdnnpCalcNeighborLists(h, ats, nl)
        call hdnnpCalcSymfunctions(h, ats, nl, sfs, together)
        call hdnnpCalcNNs(h, ats, nl, sfs, energy, force, dEdLat)
        if (.not. ats%periodic) dEdLat = 0._dp
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
        close(u)
    end subroutine dumpStages

    subroutine runMD(h, xyz, pfile, fname)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: xyz, pfile, fname
        type(atStruct) :: ats
        real(dp), allocatable :: forces(:,:), m(:,:), p(:,:), lastForces(:,:)
        real(dp), parameter :: stepsize = 0.5_dp * femtoseconds2atomic
        real(dp) :: epot, ekin
        integer :: i, u, nsteps
        character(len=64) :: arg

        call get_command_argument(5, arg)
        read(arg, *) nsteps
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
        write(u, '(A)') 'momenta'
        do i = 1, ats%nat
            write(u, '(3ES25.16E3)') p(:, i)
        end do
        close(u)
    end subroutine runMD

    subroutine runTime(h, sfile, fname)
        type(hdnnp), intent(inout) :: h
        character(len=*), intent(in) :: sfile, fname
        type(atStruct) :: ats
        real(dp) :: energy, dEdLat(3,3)
        real(dp), allocatable :: force(:,:), t(:)
        integer(8) :: c0, c1, rate
        integer :: u, k, nrep
        character(len=64) :: arg

        call get_command_argument(4, arg)
        read(arg, *) nrep
        call readStruct(sfile, ats)
        allocate(force(3, ats%nat), t(nrep))
        call hdnnpEnergyAndForces(h, ats, energy, force, dEdLat)
        do k = 1, nrep
            call system_clock(c0, rate)
            call hdnnpEnergyAndForces(h, ats, energy, force, dEdLat)
            call system_clock(c1)
            t(k) = real(c1 - c0, dp) / real(rate, dp)
        end do
        open(newunit=u, file=fname, action='write', status='replace')
        write(u, '(A,I10)') 'nat ', ats%nat
        write(u, '(A,ES25.16E3)') 'energy ', energy
        do k = 1, nrep
            write(u, '(A,ES25.16E3)') 'seconds ', t(k)
        end do
        close(u)
    end subroutine runTime

end program refdump
