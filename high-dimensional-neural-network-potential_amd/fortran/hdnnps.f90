! Fortran host side of the MI355X HDNNP engine.
!
! These modules carry the public names of the reference's Fortran interface
! so that a driver written against it (e.g. src/potentials/hdnnpExample.f90)
! compiles against this package:
!   precision        -- dp kind                         (src/precision.f90)
!   constants        -- units, masses, element symbols  (src/constants/constants.f90)
!   atomicStructure  -- type atStruct, as_readXYZ, as_writeXYZ,
!                       as_readFhiaims, as_readAscii, as_readPOSCAR,
!                       as_readData, as_repeatUnitCell, dEdLat2Stress,
!                       stress2dEdLat, center          (src/atoms/atomicStructure.f90)
!   hdnnps           -- type hdnnp, parseRuNNerInput, hdnnpEnergyAndForces
!                                                       (src/potentials/hdnnps.f90)
! The evaluation runs on the GPU through the C ABI of libhdnnp_amd.so
! (include/hdnnp_capi.h), bound here with iso_c_binding. Errors end the
! program with `error stop` and the library's message, as the reference's
! print + stop do.

module precision
    use iso_c_binding
    implicit none
    integer, parameter :: dp = c_double
end module precision

module constants
    use precision
    implicit none
    real(dp), parameter :: ang2bohr = 1.8897259886_dp
    real(dp), parameter :: ev2ha = 0.03674932539796232_dp
    real(dp), parameter :: kBoltzmann = 3.166811965e-6_dp
    real(dp), parameter :: massUnit2ElectronMass = 1822.888513874368_dp
    real(dp), parameter :: femtoseconds2atomic = 41.34137333519408_dp
    integer, parameter :: maxElemNum = 118
    character(len=2), parameter :: elemSymbols(118) = [ &
        ' H', 'He', 'Li', 'Be', ' B', ' C', ' N', ' O', ' F', 'Ne', 'Na', 'Mg', 'Al', 'Si', ' P', ' S', &
        'Cl', 'Ar', ' K', 'Ca', 'Sc', 'Ti', ' V', 'Cr', 'Mn', 'Fe', 'Co', 'Ni', 'Cu', 'Zn', 'Ga', 'Ge', &
        'As', 'Se', 'Br', 'Kr', 'Rb', 'Sr', ' Y', 'Zr', 'Nb', 'Mo', 'Tc', 'Ru', 'Rh', 'Pd', 'Ag', 'Cd', &
        'In', 'Sn', 'Sb', 'Te', ' I', 'Xe', 'Cs', 'Ba', 'La', 'Ce', 'Pr', 'Nd', 'Pm', 'Sm', 'Eu', 'Gd', &
        'Tb', 'Dy', 'Ho', 'Er', 'Tm', 'Yb', 'Lu', 'Hf', 'Ta', ' W', 'Re', 'Os', 'Ir', 'Pt', 'Au', 'Hg', &
        'Tl', 'Pb', 'Bi', 'Po', 'At', 'Rn', 'Fr', 'Ra', 'Ac', 'Th', 'Pa', ' U', 'Np', 'Pu', 'Am', 'Cm', &
        'Bk', 'Cf', 'Es', 'Fm', 'Md', 'No', 'Lr', 'Rf', 'Db', 'Sg', 'Bh', 'Hs', 'Mt', 'Ds', 'Rg', 'Cn', &
        'Nh', 'Fl', 'Mc', 'Lv', 'Ts', 'Og']
    ! IUPAC 2019 standard atomic weights of H..Ar, and I, Pb (amu); other
    ! elements map to -1
    real(dp), parameter :: lightMasses(18) = [1.008_dp, 4.002_dp, 6.94_dp, 9.012_dp, 10.81_dp, &
        12.011_dp, 14.007_dp, 15.999_dp, 18.998_dp, 20.1797_dp, 22.989_dp, 24.305_dp, 26.981_dp, &
        28.085_dp, 30.973_dp, 32.06_dp, 35.45_dp, 39.948_dp]
contains
    elemental function getAtomicMass(n) result(m)
        integer, intent(in) :: n
        real(dp) :: m
        m = -1._dp
        if (n >= 1 .and. n <= 18) m = lightMasses(n)
        if (n == 53) m = 126.904_dp
        if (n == 82) m = 207.2_dp
    end function getAtomicMass

    elemental function elemSymToNum(sym) result(n)
        character(len=*), intent(in) :: sym
        integer :: n, i
        n = -1
        do i = 1, 118
            if (trim(adjustl(sym)) == trim(adjustl(elemSymbols(i)))) then
                n = i
                return
            end if
        end do
    end function elemSymToNum

    elemental function elemNumToSym(n) result(sym)
        integer, intent(in) :: n
        character(len=2) :: sym
        sym = 'XX'
        if (n >= 1 .and. n <= 118) sym = elemSymbols(n)
    end function elemNumToSym
end module constants

module hdnnp_capi
    use iso_c_binding
    implicit none
    integer(c_int), parameter :: HDNNP_FP64 = 0, HDNNP_MIXED = 1
    ! hdnnp_model_info of include/hdnnp_capi.h, field for field
    type, bind(C) :: hdnnp_model_info
        integer(c_int) :: nelements, Z(16), nsf(16), nlayers, nodes(8)
        character(kind=c_char) :: act(8)
        integer(c_int) :: scalesf, centersf
        real(c_double) :: maxrc, atom_energy(16), emin, emax
        integer(c_int) :: cutoff_type, scale_range
        real(c_double) :: scale_min, scale_max
    end type hdnnp_model_info
    interface
        integer(c_int) function hdnnp_create(dir, precision, ctx) bind(C, name='hdnnp_create')
            import :: c_int, c_char, c_ptr
            character(kind=c_char), dimension(*), intent(in) :: dir
            integer(c_int), value :: precision
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function hdnnp_evaluate(ctx, nat, Z, pos, periodic, lat, energy, force, dEdLat) &
                bind(C, name='hdnnp_evaluate')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nat, periodic
            integer(c_int), intent(in) :: Z(*)
            real(c_double), intent(in) :: pos(*), lat(*)
            real(c_double), intent(out) :: energy
            real(c_double), intent(out) :: force(*)
            type(c_ptr), value :: dEdLat
        end function
        integer(c_int) function hdnnp_dEdLat_to_stress(lat, dEdLat, stress) bind(C, name='hdnnp_dEdLat_to_stress')
            import :: c_int, c_double
            real(c_double), intent(in) :: lat(*), dEdLat(*)
            real(c_double), intent(out) :: stress(*)
        end function
        type(c_ptr) function hdnnp_last_error(ctx) bind(C, name='hdnnp_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        subroutine hdnnp_destroy(ctx) bind(C, name='hdnnp_destroy')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        integer(c_int) function hdnnp_create_committee(n, dirs, precision, ctx) &
                bind(C, name='hdnnp_create_committee')
            import :: c_int, c_ptr
            integer(c_int), value :: n, precision
            type(c_ptr), intent(in) :: dirs(*)
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function hdnnp_last_committee(ctx, nstruct, std_dev, member_energies) &
                bind(C, name='hdnnp_last_committee')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nstruct
            real(c_double), intent(out) :: std_dev
            type(c_ptr), value :: member_energies
        end function
        integer(c_int) function hdnnp_get_model_info(ctx, info) bind(C, name='hdnnp_get_model_info')
            import :: c_int, c_ptr, hdnnp_model_info
            type(c_ptr), value :: ctx
            type(hdnnp_model_info), intent(out) :: info
        end function
        integer(c_int) function hdnnp_get_sf_table(ctx, e, cap_rows, rows) bind(C, name='hdnnp_get_sf_table')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: e, cap_rows
            real(c_double), intent(out) :: rows(*)
        end function
        integer(c_int) function hdnnp_stage_neighbor_lists(ctx, nat, Z, pos, periodic, lat, stage, nneis) &
                bind(C, name='hdnnp_stage_neighbor_lists')
            import :: c_int, c_double, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: nat, periodic
            integer(c_int), intent(in) :: Z(*)
            real(c_double), intent(in) :: pos(*), lat(*)
            integer(c_int64_t), intent(out) :: stage
            integer(c_int), intent(out) :: nneis(*)
        end function
        integer(c_int) function hdnnp_stage_get_neighbor_lists(ctx, stage, ineis, neis) &
                bind(C, name='hdnnp_stage_get_neighbor_lists')
            import :: c_int, c_double, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: stage
            integer(c_int), intent(out) :: ineis(*)
            real(c_double), intent(out) :: neis(*)
        end function
        integer(c_int) function hdnnp_stage_symfunctions(ctx, stage, nat, Z, pos, periodic, lat, together, ld, G, dsf) &
                bind(C, name='hdnnp_stage_symfunctions')
            import :: c_int, c_double, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(inout) :: stage
            integer(c_int), value :: nat, periodic, together, ld
            integer(c_int), intent(in) :: Z(*)
            real(c_double), intent(in) :: pos(*), lat(*)
            real(c_double), intent(out) :: G(*), dsf(*)
        end function
        integer(c_int) function hdnnp_stage_nns(net, sfctx, stage, nat, Z, pos, periodic, lat, ld, G, energy, &
                force, dEdLat) bind(C, name='hdnnp_stage_nns')
            import :: c_int, c_double, c_ptr, c_int64_t
            type(c_ptr), value :: net, sfctx
            integer(c_int64_t), intent(inout) :: stage
            integer(c_int), value :: nat, periodic, ld
            integer(c_int), intent(in) :: Z(*)
            real(c_double), intent(in) :: pos(*), lat(*), G(*)
            real(c_double), intent(out) :: energy, force(*)
            type(c_ptr), value :: dEdLat
        end function
    end interface
contains
    function c_message(p) result(msg)
        type(c_ptr), intent(in) :: p
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        integer :: n
        interface
            function strlen(q) bind(C, name='strlen')
                import :: c_ptr, c_size_t
                type(c_ptr), value :: q
                integer(c_size_t) :: strlen
            end function
        end interface
        n = int(strlen(p))
        call c_f_pointer(p, s, [n])
        allocate(character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function c_message
end module hdnnp_capi

module atomicStructure
    use precision
    use constants
    use hdnnp_capi
    implicit none

    ! same components as the reference's atStruct (atomicStructure.f90:17-28)
    type atStruct
        integer :: nat
        logical :: periodic = .false.
        real(dp) :: lat(3,3)
        real(dp), allocatable :: ats(:,:), f(:,:), q(:)
        integer, allocatable :: el(:)
        real(dp), allocatable :: nValenceElectrons(:)
        real(dp), allocatable :: rnuc(:)
        real(dp) :: qtot, energy
    end type atStruct

contains

    ! as_readXYZ (atomicStructure.f90:143-166): coordinates in bohr, nonperiodic
    subroutine as_readXYZ(filename, ats)
        character(len=*), intent(in) :: filename
        type(atStruct), intent(out) :: ats
        integer :: i, u, stat
        character(len=2) :: el
        open(newunit=u, file=filename, iostat=stat, status='old', action='read')
        if (stat /= 0) then
            print*, 'Error: Could not open file (' // filename // ')'
            error stop 1
        end if
        read(u, *) ats%nat
        read(u, *)
        allocate(ats%ats(3, ats%nat), ats%el(ats%nat))
        do i = 1, ats%nat
            read(u, *) el, ats%ats(:, i)
            ats%el(i) = elemSymToNum(el)
        end do
        ats%periodic = .false.
        ats%lat = 0._dp
        close(u)
    end subroutine as_readXYZ

    subroutine as_writeXYZ(filename, ats)
        character(len=*), intent(in) :: filename
        type(atStruct), intent(in) :: ats
        integer :: i, u
        open(newunit=u, file=filename, action='write', status='replace')
        write(u, *) ats%nat
        write(u, *)
        do i = 1, ats%nat
            write(u, *) elemNumToSym(ats%el(i)), ats%ats(:, i)
        end do
        close(u)
    end subroutine as_writeXYZ

    ! file cannot be opened: the reference's message, then stop
    subroutine openOrStop(filename, u)
        character(len=*), intent(in) :: filename
        integer, intent(out) :: u
        integer :: stat
        open(newunit=u, file=filename, iostat=stat, status='old', action='read')
        if (stat /= 0) then
            print*, 'Error: Could not open file (' // filename // ')'
            error stop 1
        end if
    end subroutine openOrStop

    ! copies of a (rows x cols) and of a vector with room for cap columns / entries
    subroutine growCols(a, cap)
        real(dp), allocatable, intent(inout) :: a(:,:)
        integer, intent(in) :: cap
        real(dp), allocatable :: t(:,:)
        allocate(t(size(a, 1), cap))
        t(:, 1:size(a, 2)) = a
        call move_alloc(t, a)
    end subroutine growCols

    subroutine growReal(a, cap)
        real(dp), allocatable, intent(inout) :: a(:)
        integer, intent(in) :: cap
        real(dp), allocatable :: t(:)
        allocate(t(cap))
        t(1:size(a)) = a
        call move_alloc(t, a)
    end subroutine growReal

    subroutine growInt(a, cap)
        integer, allocatable, intent(inout) :: a(:)
        integer, intent(in) :: cap
        integer, allocatable :: t(:)
        allocate(t(cap))
        t(1:size(a)) = a
        call move_alloc(t, a)
    end subroutine growInt

    ! room for at least n atoms in positions p(3,:), elements e and, when
    ! given, charges q and forces f(3,:)
    subroutine growAtoms(n, p, e, q, f)
        integer, intent(in) :: n
        real(dp), allocatable, intent(inout) :: p(:,:)
        integer, allocatable, intent(inout) :: e(:)
        real(dp), allocatable, intent(inout), optional :: q(:), f(:,:)
        integer :: cap
        if (n <= size(e)) return
        cap = max(2 * size(e), n, 16)
        call growCols(p, cap)
        call growInt(e, cap)
        if (present(q)) call growReal(q, cap)
        if (present(f)) call growCols(f, cap)
    end subroutine growAtoms

    ! as_readFhiaims (atomicStructure.f90:168-219): FHI-aims geometry in
    ! Angstrom; lines beginning with 'atom' in column 1 and 'lattice_vector'
    ! lines; periodic when a lattice vector is given
    subroutine as_readFhiaims(filename, ats)
        character(len=*), intent(in) :: filename
        type(atStruct), intent(out) :: ats
        character(len=200) :: line, tag
        character(len=2) :: el
        real(dp) :: lat(3,3), v(3)
        real(dp), allocatable :: p(:,:)
        integer, allocatable :: e(:)
        integer :: u, stat, n, nlat
        call openOrStop(filename, u)
        allocate(p(3, 0), e(0))
        n = 0
        nlat = 0
        lat = 0._dp
        do
            read(u, '(A)', iostat=stat) line
            if (stat /= 0) exit
            if (line(1:4) == 'atom') then
                read(line, *) tag, v, el
                n = n + 1
                call growAtoms(n, p, e)
                p(:, n) = v
                e(n) = elemSymToNum(el)
            else if (line(1:14) == 'lattice_vector') then
                if (nlat == 3) error stop 'ERROR: more than 3 lattice vectors'
                nlat = nlat + 1
                read(line, *) tag, lat(:, nlat)
            end if
        end do
        close(u)
        ats%nat = n
        allocate(ats%ats(3, n), ats%el(n))
        ats%ats = p(:, 1:n) * ang2bohr
        ats%el = e(1:n)
        ats%periodic = nlat > 0
        ats%lat = 0._dp
        if (ats%periodic) ats%lat = lat * ang2bohr
    end subroutine as_readFhiaims

    ! as_readAscii (atomicStructure.f90:245-285): v_sim ascii, periodic; the
    ! lattice is lower triangular (dxx / dyx dyy on line 2, dzx dzy dzz on
    ! line 3); Angstrom when line 1 says 'angstroem' or isAngstrom is set
    subroutine as_readAscii(filename, ats, isAngstrom)
        character(len=*), intent(in) :: filename
        type(atStruct), intent(out) :: ats
        logical, intent(in), optional :: isAngstrom
        character(len=200) :: line, unit
        character(len=2) :: el
        integer :: u, stat, i
        logical :: angs
        call openOrStop(filename, u)
        read(u, '(A)') line
        unit = ''
        read(line, *, iostat=stat) ats%nat, unit
        if (stat /= 0) read(line, *) ats%nat
        angs = unit(1:9) == 'angstroem'
        if (present(isAngstrom)) angs = angs .or. isAngstrom
        ats%periodic = .true.
        allocate(ats%ats(3, ats%nat), ats%el(ats%nat))
        ats%lat = 0._dp
        read(u, *) ats%lat(1:1, 1), ats%lat(1:2, 2)
        read(u, *) ats%lat(1:3, 3)
        do i = 1, ats%nat
            read(u, *) ats%ats(:, i), el
            ats%el(i) = elemSymToNum(el)
        end do
        close(u)
        if (angs) then
            ats%ats = ats%ats * ang2bohr
            ats%lat = ats%lat * ang2bohr
        end if
    end subroutine as_readAscii

    ! as_readPOSCAR (atomicStructure.f90:427-475): VASP 5 with element names
    ! and Direct coordinates; the scale factor is read but not applied, as in
    ! the reference
    subroutine as_readPOSCAR(fname, ats)
        character(len=*), intent(in) :: fname
        type(atStruct), intent(out) :: ats
        character(len=256) :: line
        character(len=2), allocatable :: syms(:)
        integer, allocatable :: cnt(:)
        real(dp) :: scale, s(3)
        integer :: u, i, j, c, n, a
        call openOrStop(fname, u)
        read(u, *)
        read(u, *) scale
        do i = 1, 3
            read(u, *) ats%lat(:, i)
        end do
        ats%lat = ats%lat * ang2bohr
        ats%periodic = .true.
        read(u, '(A)') line
        n = 0                                  ! blank-separated words
        do i = 1, len_trim(line)
            if (line(i:i) /= ' ' .and. (i == 1 .or. line(max(i - 1, 1):max(i - 1, 1)) == ' ')) n = n + 1
        end do
        allocate(syms(n), cnt(n))
        read(line, *) syms
        read(u, *) cnt
        read(u, '(A)') line
        do i = 1, len(line)
            c = ichar(line(i:i))
            if (c >= 65 .and. c <= 90) line(i:i) = char(c + 32)
        end do
        if (trim(adjustl(line)) /= 'direct') then
            print*, line
            error stop 'ERROR: POSCAR without DIRECT not implemented'
        end if
        ats%nat = sum(cnt)
        allocate(ats%el(ats%nat), ats%ats(3, ats%nat))
        c = 0
        do i = 1, n
            do j = 1, cnt(i)
                c = c + 1
                ats%el(c) = elemSymToNum(syms(i))
            end do
        end do
        do i = 1, ats%nat                      ! lat . s (matMulVec3, linalg.f90:192-201)
            read(u, *) s
            do a = 1, 3
                ats%ats(a, i) = (ats%lat(a, 1) * s(1) + ats%lat(a, 2) * s(2)) + ats%lat(a, 3) * s(3)
            end do
        end do
        close(u)
    end subroutine as_readPOSCAR

    ! as_readData (atomicStructure.f90:354-425): RuNNer input.data blocks
    ! (begin / lattice / atom x y z el q e fx fy fz / energy / charge / end),
    ! bohr and Hartree; periodic when exactly three lattice lines are given
    subroutine as_readData(fname, n, ats)
        character(len=*), intent(in) :: fname
        integer, intent(out) :: n
        type(atStruct), allocatable, intent(out) :: ats(:)
        type(atStruct), allocatable :: tmp(:)
        character(len=512) :: line
        character(len=64) :: tag
        character(len=2) :: el
        real(dp) :: lat(3,3), v(3), f(3), q, e, energy, qtot
        real(dp), allocatable :: p(:,:), fs(:,:), qs(:)
        integer, allocatable :: es(:)
        integer :: u, stat, na, nlat, k
        logical :: inside
        call openOrStop(fname, u)
        allocate(ats(0), p(3, 0), es(0), qs(0), fs(3, 0))
        n = 0
        inside = .false.
        do
            read(u, '(A)', iostat=stat) line
            if (stat /= 0) exit
            line = adjustl(line)
            if (line(1:5) == 'begin') then
                inside = .true.
                na = 0
                nlat = 0
                lat = 0._dp
                energy = 0._dp
                qtot = 0._dp
            end if
            if (.not. inside) cycle
            if (line(1:7) == 'lattice') then
                if (nlat == 3) error stop 'ERROR: more than 3 lattice lines in a structure'
                nlat = nlat + 1
                read(line, *) tag, lat(:, nlat)
            end if
            if (line(1:4) == 'atom') then
                read(line, *) tag, v, el, q, e, f
                na = na + 1
                call growAtoms(na, p, es, qs, fs)
                p(:, na) = v
                es(na) = elemSymToNum(el)
                qs(na) = q
                fs(:, na) = f
            end if
            if (line(1:6) == 'energy') read(line, *) tag, energy
            if (line(1:6) == 'charge') read(line, *) tag, qtot
            if (line(1:3) == 'end') then
                allocate(tmp(n + 1))
                do k = 1, n
                    tmp(k) = ats(k)
                end do
                associate (s => tmp(n + 1))
                    s%nat = na
                    s%periodic = nlat == 3
                    s%lat = 0._dp
                    if (s%periodic) s%lat = lat
                    allocate(s%ats(3, na), s%el(na), s%q(na), s%f(3, na))
                    s%ats = p(:, 1:na)
                    s%el = es(1:na)
                    s%q = qs(1:na)
                    s%f = fs(:, 1:na)
                    s%energy = energy
                    s%qtot = qtot
                end associate
                call move_alloc(tmp, ats)
                n = n + 1
                inside = .false.
            end if
        end do
        close(u)
    end subroutine as_readData

    ! as_repeatUnitCell (atomicStructure.f90:636-671): x, y, z loops, atoms innermost
    subroutine as_repeatUnitCell(ats, n)
        type(atStruct), intent(inout) :: ats
        integer, intent(in) :: n(3)
        real(dp), allocatable :: p(:,:)
        integer, allocatable :: e(:)
        integer :: x, y, z, i, c
        allocate(p(3, ats%nat * product(n)), e(ats%nat * product(n)))
        c = 0
        do x = 1, n(1)
            do y = 1, n(2)
                do z = 1, n(3)
                    do i = 1, ats%nat
                        e(c * ats%nat + i) = ats%el(i)
                        p(:, c * ats%nat + i) = ats%ats(:, i) + (x - 1) * ats%lat(:, 1) + &
                            (y - 1) * ats%lat(:, 2) + (z - 1) * ats%lat(:, 3)
                    end do
                    c = c + 1
                end do
            end do
        end do
        ats%lat(:, 1) = ats%lat(:, 1) * n(1)
        ats%lat(:, 2) = ats%lat(:, 2) * n(2)
        ats%lat(:, 3) = ats%lat(:, 3) * n(3)
        call move_alloc(p, ats%ats)
        call move_alloc(e, ats%el)
        ats%nat = ats%nat * product(n)
    end subroutine as_repeatUnitCell

    ! dEdLat2Stress (atomicStructure.f90:679-686)
    subroutine dEdLat2Stress(lat, dEdLat, stress)
        real(dp), intent(in) :: lat(3,3)
        real(dp), intent(out) :: stress(3,3)
        real(dp), intent(in) :: dEdLat(3,3)
        if (hdnnp_dEdLat_to_stress(lat, dEdLat, stress) /= 0) error stop 'dEdLat2Stress failed'
    end subroutine dEdLat2Stress

    ! stress2dEdLat (atomicStructure.f90:688-697): dEdLat = -det(lat) stress lat^-T
    subroutine stress2dEdLat(lat, stress, dEdLat)
        real(dp), intent(in) :: lat(3,3), stress(3,3)
        real(dp), intent(out) :: dEdLat(3,3)
        real(dp) :: inv(3,3), det
        det = lat(1,1) * lat(2,2) * lat(3,3) + lat(1,2) * lat(2,3) * lat(3,1) &
            + lat(1,3) * lat(2,1) * lat(3,2) - lat(1,1) * lat(2,3) * lat(3,2) &
            - lat(1,2) * lat(2,1) * lat(3,3) - lat(1,3) * lat(2,2) * lat(3,1)
        inv(1,1) = lat(2,2) * lat(3,3) - lat(2,3) * lat(3,2)
        inv(1,2) = lat(1,3) * lat(3,2) - lat(1,2) * lat(3,3)
        inv(1,3) = lat(1,2) * lat(2,3) - lat(1,3) * lat(2,2)
        inv(2,1) = lat(2,3) * lat(3,1) - lat(2,1) * lat(3,3)
        inv(2,2) = lat(1,1) * lat(3,3) - lat(1,3) * lat(3,1)
        inv(2,3) = lat(1,3) * lat(2,1) - lat(1,1) * lat(2,3)
        inv(3,1) = lat(2,1) * lat(3,2) - lat(2,2) * lat(3,1)
        inv(3,2) = lat(1,2) * lat(3,1) - lat(1,1) * lat(3,2)
        inv(3,3) = lat(1,1) * lat(2,2) - lat(1,2) * lat(2,1)
        inv = inv / det
        dEdLat = -1._dp * det * matmul(stress, transpose(inv))
    end subroutine stress2dEdLat

    ! center (pointsets.f90:245-256): remove the mean
    subroutine center(nat, ats)
        integer, intent(in) :: nat
        real(dp), intent(inout), dimension(:, :) :: ats
        real(dp) :: cm(3)
        integer :: i
        cm = sum(ats, 2) / nat
        do i = 1, nat
            ats(:, i) = ats(:, i) - cm(:)
        end do
    end subroutine center
end module atomicStructure

! type neighborList of the reference (src/potentials/neighborLists.f90:10-16):
! slot 1 is the atom itself; neis holds the positions of the images, ineis
! their atom numbers; arrays grow from 16 by doubling (maxNeis).
module neighborLists
    use precision
    implicit none
    type neighborList
        integer :: nNeis = 0
        integer :: maxNeis = 0
        real(dp), allocatable :: neis(:,:)
        integer, allocatable :: ineis(:)
        ! not in the reference: the library stage the list came from
        integer(c_int64_t) :: stage = 0
    end type neighborList
end module neighborLists

! types of the reference's symmetry functions (src/potentials/
! symmetryFunctions.f90:7-33): the parsed functions of an element
! (symFunctionList) and the values and gradients of one atom
! (symFunctionContainer; dsf(:, n) = dG/dR of list slot n).
module symmetryFunctions
    use precision
    implicit none
    type symFunction
        integer :: name
        integer :: sfEls(2)
        real(dp) :: rcut
        real(dp) :: eta
        real(dp) :: lambda
        real(dp) :: zeta
        real(dp) :: Rs
        real(dp) :: sfmin, sfmax, sfmean
    end type symFunction

    type symFunctionList
        integer :: nsf = 0
        type(symFunction), allocatable :: sfs(:)
    end type symFunctionList

    type symFunctionValue
        integer :: nNeis
        real(dp) :: sf
        real(dp), allocatable :: dsf(:,:)
    end type symFunctionValue

    type symFunctionContainer
        integer :: nsf = 0
        type(symFunctionValue), allocatable :: sfs(:)
        ! not in the reference: the context whose lists and symmetry functions
        ! these are, and its stage
        type(c_ptr) :: ctx = c_null_ptr
        integer(c_int64_t) :: stage = 0
    end type symFunctionContainer
end module symmetryFunctions

module hdnnps
    use precision
    use atomicStructure
    use constants
    use hdnnp_capi
    use neighborLists
    use symmetryFunctions
    implicit none

    ! the reference's type(hdnnp) (hdnnps.f90:13-21). The parsed tables live
    ! in the library context behind `ctx`; sfLists holds a copy of the sorted
    ! symmetry functions (the networks, nnList, stay in the library).
    type hdnnp
        type(c_ptr) :: ctx = c_null_ptr
        type(symFunctionList) :: sfLists(maxElemNum)
        real(dp) :: maxRc = 0._dp
        real(dp) :: emin = 0._dp, emax = 0._dp
        real(dp) :: atomenergies(maxElemNum) = 0._dp
        logical :: scalesf = .false.
        logical :: centersf = .false.
    end type hdnnp

contains

    ! initHdnnp (hdnnps.f90:25-32): an empty handle
    subroutine initHdnnp(handle)
        type(hdnnp), intent(out) :: handle
        handle%atomenergies = 0._dp
        handle%scalesf = .false.
        handle%centersf = .false.
    end subroutine initHdnnp

    ! parseRuNNerInput (hdnnps.f90:34-233). `path` is the directory prefix of
    ! input.nn, scaling.data and weights.ZZZ.data. `precision` (optional,
    ! not in the reference) selects HDNNP_FP64 (default) or HDNNP_MIXED.
    subroutine parseRuNNerInput(handle, path, precision)
        type(hdnnp), intent(inout) :: handle
        character(len=*) :: path
        integer, intent(in), optional :: precision
        integer(c_int) :: prec, rc
        prec = HDNNP_FP64
        if (present(precision)) prec = int(precision, c_int)
        if (c_associated(handle%ctx)) call hdnnp_destroy(handle%ctx)
        rc = hdnnp_create(trim(path) // c_null_char, prec, handle%ctx)
        if (rc /= 0) then
            print*, c_message(hdnnp_last_error(c_null_ptr))
            error stop 1
        end if
        call fillHandle(handle)
    end subroutine parseRuNNerInput

    ! the handle's public fields from the library context
    subroutine fillHandle(handle)
        type(hdnnp), intent(inout) :: handle
        type(hdnnp_model_info) :: info
        integer(c_int) :: rc
        integer :: e
        rc = hdnnp_get_model_info(handle%ctx, info)
        handle%maxRc = info%maxrc
        handle%emin = info%emin
        handle%emax = info%emax
        handle%scalesf = info%scalesf /= 0
        handle%centersf = info%centersf /= 0
        handle%atomenergies = 0._dp
        do e = 1, info%nelements
            handle%atomenergies(info%Z(e)) = info%atom_energy(e)
        end do
        call fillSfLists(handle, info)
    end subroutine fillHandle

    ! handle%sfLists from the library's sorted tables (rows of type, Z1, Z2,
    ! rcut, eta, Rs, lambda, zeta, min, max, mean)
    subroutine fillSfLists(handle, info)
        type(hdnnp), intent(inout) :: handle
        type(hdnnp_model_info), intent(in) :: info
        real(dp), allocatable :: rows(:,:)
        integer :: e, s, z, n
        integer(c_int) :: rc
        do z = 1, maxElemNum
            handle%sfLists(z)%nsf = 0
            if (allocated(handle%sfLists(z)%sfs)) deallocate(handle%sfLists(z)%sfs)
        end do
        do e = 1, info%nelements
            z = info%Z(e)
            n = info%nsf(e)
            allocate(rows(11, max(n, 1)), handle%sfLists(z)%sfs(n))
            rc = hdnnp_get_sf_table(handle%ctx, int(e - 1, c_int), int(n, c_int), rows)
            handle%sfLists(z)%nsf = n
            do s = 1, n
                handle%sfLists(z)%sfs(s)%name = nint(rows(1, s))
                handle%sfLists(z)%sfs(s)%sfEls = [nint(rows(2, s)), nint(rows(3, s))]
                handle%sfLists(z)%sfs(s)%rcut = rows(4, s)
                handle%sfLists(z)%sfs(s)%eta = rows(5, s)
                handle%sfLists(z)%sfs(s)%Rs = rows(6, s)
                handle%sfLists(z)%sfs(s)%lambda = rows(7, s)
                handle%sfLists(z)%sfs(s)%zeta = rows(8, s)
                handle%sfLists(z)%sfs(s)%sfmin = rows(9, s)
                handle%sfLists(z)%sfs(s)%sfmax = rows(10, s)
                handle%sfLists(z)%sfs(s)%sfmean = rows(11, s)
            end do
            deallocate(rows)
        end do
    end subroutine fillSfLists

    ! the structure as the C ABI takes it
    subroutine cStructure(ats, per, lat)
        type(atStruct), intent(in) :: ats
        integer(c_int), intent(out) :: per
        real(dp), intent(out) :: lat(3,3)
        per = 0
        lat = 0._dp
        if (ats%periodic) then
            per = 1
            lat = ats%lat
        end if
    end subroutine cStructure

    subroutine stopOnError(rc, ctx)
        integer(c_int), intent(in) :: rc
        type(c_ptr), intent(in) :: ctx
        if (rc /= 0) then
            print*, c_message(hdnnp_last_error(ctx))
            error stop 1
        end if
    end subroutine stopOnError

    ! hdnnpCalcNeighborLists (hdnnps.f90:373-388): the linked-cell lists of
    ! the (wrapped) structure within maxRc, built on the device, in the
    ! reference's layout and order
    subroutine hdnnpCalcNeighborLists(handle, ats, neiLists)
        type(hdnnp), intent(in) :: handle
        type(atStruct), intent(in) :: ats
        type(neighborList), intent(out) :: neiLists(ats%nat)
        integer(c_int) :: per, rc
        real(dp) :: lat(3,3)
        integer(c_int64_t) :: stage
        integer(c_int), allocatable :: nneis(:), ineis(:)
        real(dp), allocatable :: neis(:,:)
        integer :: i, q, cap
        call cStructure(ats, per, lat)
        allocate(nneis(max(ats%nat, 1)))
        rc = hdnnp_stage_neighbor_lists(handle%ctx, int(ats%nat, c_int), ats%el, ats%ats, per, lat, stage, nneis)
        call stopOnError(rc, handle%ctx)
        allocate(ineis(max(sum(nneis(1:ats%nat)), 1)), neis(3, max(sum(nneis(1:ats%nat)), 1)))
        rc = hdnnp_stage_get_neighbor_lists(handle%ctx, stage, ineis, neis)
        call stopOnError(rc, handle%ctx)
        q = 0
        do i = 1, ats%nat
            cap = 16                     ! the reference's capacity: 16, doubled when full
            do while (cap < nneis(i))
                cap = 2 * cap
            end do
            neiLists(i)%nNeis = nneis(i)
            neiLists(i)%maxNeis = cap
            allocate(neiLists(i)%neis(3, cap), neiLists(i)%ineis(cap))
            neiLists(i)%neis = 0._dp
            neiLists(i)%ineis = -1
            neiLists(i)%neis(:, 1:nneis(i)) = neis(:, q + 1:q + nneis(i))
            neiLists(i)%ineis(1:nneis(i)) = ineis(q + 1:q + nneis(i))
            neiLists(i)%stage = stage
            q = q + nneis(i)
        end do
    end subroutine hdnnpCalcNeighborLists

    ! hdnnpCalcSymfunctions (hdnnps.f90:390-485): centred and scaled values
    ! and dG/dR of every atom's symmetry functions. calcSfsTogether (default
    ! .true.) selects the fused pass (calcAllSymmetryFunctions) for the
    ! values; .false. the per-function kernel (symmetryFunction2/3), which
    ! gives the gradients either way.
    subroutine hdnnpCalcSymfunctions(handle, ats, neiLists, sfs, calcSfsTogether)
        type(hdnnp), intent(in) :: handle
        type(atStruct), intent(in) :: ats
        type(neighborList), intent(in) :: neiLists(ats%nat)
        type(symFunctionContainer), intent(out) :: sfs(ats%nat)
        logical, intent(in), optional :: calcSfsTogether
        integer(c_int) :: per, rc, together, ld
        real(dp) :: lat(3,3)
        integer(c_int64_t) :: stage
        real(dp), allocatable :: G(:,:), dsf(:)
        integer(c_int64_t) :: total, o
        integer :: i, s, nsf, nn
        together = 1
        if (present(calcSfsTogether)) then
            if (.not. calcSfsTogether) together = 0
        end if
        call cStructure(ats, per, lat)
        stage = 0
        if (ats%nat > 0) stage = neiLists(1)%stage
        ld = 1
        total = 0
        do i = 1, ats%nat
            nsf = handle%sfLists(ats%el(i))%nsf
            ld = max(ld, int(nsf, c_int))
            total = total + 3_c_int64_t * nsf * neiLists(i)%nNeis
        end do
        allocate(G(ld, max(ats%nat, 1)), dsf(max(total, 1_c_int64_t)))
        rc = hdnnp_stage_symfunctions(handle%ctx, stage, int(ats%nat, c_int), ats%el, ats%ats, per, lat, &
                                      together, ld, G, dsf)
        call stopOnError(rc, handle%ctx)
        o = 0
        do i = 1, ats%nat
            nsf = handle%sfLists(ats%el(i))%nsf
            nn = neiLists(i)%nNeis
            sfs(i)%nsf = nsf
            sfs(i)%ctx = handle%ctx
            sfs(i)%stage = stage
            allocate(sfs(i)%sfs(nsf))
            do s = 1, nsf
                sfs(i)%sfs(s)%nNeis = nn
                sfs(i)%sfs(s)%sf = G(s, i)
                allocate(sfs(i)%sfs(s)%dsf(3, nn))
                sfs(i)%sfs(s)%dsf = reshape(dsf(o + 1:o + 3 * nn), [3, nn])
                o = o + 3 * nn
            end do
        end do
    end subroutine hdnnpCalcSymfunctions

    ! hdnnpCalcNNs (hdnnps.f90:487-562): this handle's networks and atom
    ! energies on the symmetry functions in sfs; forces and dEdLat contract
    ! dE/dG with the gradients of the symmetry functions of the context that
    ! computed sfs, recomputed on the device from that context's lists (the
    ! stage neiLists came from), so neiLists and the dsf in sfs are not read
    subroutine hdnnpCalcNNs(handle, ats, neiLists, sfs, energy, force, dEdLat)
        type(hdnnp), intent(inout) :: handle
        type(atStruct), intent(in) :: ats
        type(neighborList), intent(in) :: neiLists(ats%nat)
        type(symFunctionContainer), intent(in) :: sfs(ats%nat)
        real(dp), intent(out) :: energy
        real(dp), intent(out) :: force(3, ats%nat)
        real(dp), intent(out), optional, target :: dEdLat(3,3)
        integer(c_int) :: per, rc, ld
        real(dp) :: lat(3,3)
        integer(c_int64_t) :: stage
        type(c_ptr) :: sfctx, pl
        real(dp), allocatable :: G(:,:)
        integer :: i, s
        call cStructure(ats, per, lat)
        ld = 1
        do i = 1, ats%nat
            ld = max(ld, int(sfs(i)%nsf, c_int))
        end do
        allocate(G(ld, max(ats%nat, 1)))
        G = 0._dp
        do i = 1, ats%nat
            do s = 1, sfs(i)%nsf
                G(s, i) = sfs(i)%sfs(s)%sf
            end do
        end do
        stage = 0
        sfctx = c_null_ptr
        if (ats%nat > 0) then
            stage = sfs(1)%stage
            sfctx = sfs(1)%ctx
        end if
        if (.not. c_associated(sfctx)) sfctx = handle%ctx
        pl = c_null_ptr
        if (present(dEdLat)) pl = c_loc(dEdLat)
        rc = hdnnp_stage_nns(handle%ctx, sfctx, stage, int(ats%nat, c_int), ats%el, ats%ats, per, lat, ld, G, &
                             energy, force, pl)
        call stopOnError(rc, handle%ctx)
    end subroutine hdnnpCalcNNs

    ! hdnnpEnergyAndForces (hdnnps.f90:333-371)
    subroutine hdnnpEnergyAndForces(handle, ats, energy, force, dEdLat)
        type(hdnnp), intent(inout) :: handle
        type(atStruct), intent(inout) :: ats
        real(dp), intent(out) :: energy
        real(dp), intent(out) :: force(3, ats%nat)
        real(dp), intent(out), optional, target :: dEdLat(3,3)
        type(c_ptr) :: pl
        integer(c_int) :: rc, per
        real(dp) :: lat(3,3)
        pl = c_null_ptr
        if (present(dEdLat)) pl = c_loc(dEdLat)
        per = 0
        lat = 0._dp
        if (ats%periodic) then
            per = 1
            lat = ats%lat
        end if
        rc = hdnnp_evaluate(handle%ctx, int(ats%nat, c_int), ats%el, ats%ats, per, lat, energy, force, pl)
        if (rc /= 0) then
            print*, c_message(hdnnp_last_error(handle%ctx))
            error stop 1
        end if
    end subroutine hdnnpEnergyAndForces

    subroutine hdnnpDestroy(handle)
        type(hdnnp), intent(inout) :: handle
        if (c_associated(handle%ctx)) call hdnnp_destroy(handle%ctx)
        handle%ctx = c_null_ptr
    end subroutine hdnnpDestroy

end module hdnnps


! Committee of potentials with the reference's interface
! (src/potentials/globalHdnnps.f90). The members are one library context: the
! first potential's neighbour lists and symmetry functions, every member's
! networks; results are the committee means.
module globalHdnnps
    use precision
    use atomicStructure
    use hdnnp_capi
    use hdnnps
    implicit none

    integer :: nGlobalHdnnps = 0
    type(hdnnp) :: globalCommittee

contains

    ! initGlobalHdnnpsFromInputFile (globalHdnnps.f90:13-29): nnps.txt holds
    ! the number of potentials, then one parameter directory per line.
    ! `precision` (optional, not in the reference) as in parseRuNNerInput.
    subroutine initGlobalHdnnpsFromInputFile(precision, file)
        integer, intent(in), optional :: precision
        character(len=*), intent(in), optional :: file
        integer :: nnunit, i, k, n
        character(len=512) :: txt
        character(kind=c_char), allocatable, target :: names(:,:)
        type(c_ptr), allocatable :: ptrs(:)
        integer(c_int) :: prec, rc

        prec = HDNNP_FP64
        if (present(precision)) prec = int(precision, c_int)
        if (present(file)) then
            open(newunit=nnunit, file=file, action='read', status='old')
        else
            open(newunit=nnunit, file='nnps.txt', action='read', status='old')
        end if
        read(nnunit, *) n
        allocate(names(len(txt) + 1, n), ptrs(n))
        do i = 1, n
            read(nnunit, '(A)') txt
            txt = adjustl(txt)
            do k = 1, len_trim(txt)
                names(k, i) = txt(k:k)
            end do
            names(len_trim(txt) + 1, i) = c_null_char
            ptrs(i) = c_loc(names(1, i))
        end do
        close(nnunit)
        if (c_associated(globalCommittee%ctx)) call hdnnp_destroy(globalCommittee%ctx)
        rc = hdnnp_create_committee(int(n, c_int), ptrs, prec, globalCommittee%ctx)
        if (rc /= 0) then
            print*, c_message(hdnnp_last_error(c_null_ptr))
            error stop 1
        end if
        call fillHandle(globalCommittee)
        nGlobalHdnnps = n
    end subroutine initGlobalHdnnpsFromInputFile

    ! globalHdnnpEnergyAndForces (globalHdnnps.f90:31-70): committee mean
    ! energy, forces and dEdLat; std = population standard deviation of the
    ! member energies
    subroutine globalHdnnpEnergyAndForces(ats, energy, forces, dEdLat, std)
        type(atStruct), intent(inout) :: ats
        real(dp), intent(out) :: forces(3, ats%nat), energy
        real(dp), intent(out), optional :: std, dEdLat(3,3)
        real(dp) :: sd
        integer(c_int) :: rc
        if (present(dEdLat)) then
            call hdnnpEnergyAndForces(globalCommittee, ats, energy, forces, dEdLat)
        else
            call hdnnpEnergyAndForces(globalCommittee, ats, energy, forces)
        end if
        if (present(std)) then
            rc = hdnnp_last_committee(globalCommittee%ctx, 1_c_int, sd, c_null_ptr)
            std = sd
        end if
    end subroutine globalHdnnpEnergyAndForces

end module globalHdnnps
