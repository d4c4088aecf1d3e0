! Fortran host side of the MI355X HDNNP engine.
!
! These modules carry the public names of the reference's Fortran interface
! so that a driver written against it (e.g. src/potentials/hdnnpExample.f90)
! compiles against this package:
!   precision        -- dp kind                         (src/precision.f90)
!   constants        -- units, masses, element symbols  (src/constants/constants.f90)
!   atomicStructure  -- type atStruct, as_readXYZ, as_writeXYZ,
!                       as_repeatUnitCell, dEdLat2Stress, stress2dEdLat,
!                       center                         (src/atoms/atomicStructure.f90)
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
    type, bind(C) :: hdnnp_model_info
        integer(c_int) :: nelements, Z(16), nsf(16), nlayers, nodes(8)
        character(kind=c_char) :: act(8)
        integer(c_int) :: scalesf, centersf
        real(c_double) :: maxrc, atom_energy(16), emin, emax
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

module hdnnps
    use precision
    use atomicStructure
    use constants
    use hdnnp_capi
    implicit none

    ! the reference's type(hdnnp) (hdnnps.f90:13-21) holds the parsed tables;
    ! here they live in the library context behind `ctx`
    type hdnnp
        type(c_ptr) :: ctx = c_null_ptr
        real(dp) :: maxRc = 0._dp
        real(dp) :: emin = 0._dp, emax = 0._dp
        real(dp) :: atomenergies(maxElemNum) = 0._dp
        logical :: scalesf = .false.
        logical :: centersf = .false.
    end type hdnnp

contains

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
    end subroutine fillHandle

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
