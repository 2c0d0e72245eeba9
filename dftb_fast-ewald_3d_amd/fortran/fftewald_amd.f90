!> Fortran side of the MI355X periodic Coulomb engine: iso_c_binding interfaces to the C ABI
!> (include/fftewald_amd.h) and, on top of them, the reference library's own entry points with
!> their exact Fortran signatures (src/dftbp/extlibs/fftewald/fft/ewald_fft_3d.f90:36-76, 142-150,
!> 197-210), so that DFTB+'s `use ewald_fft_3d, only: ewald_potential_force` (coulomb.F90:25) can
!> be pointed at this module instead.
!>
!> The resident engine (TEwaldEngine) is what TCoulomb keeps between SCC iterations: created
!> once, given the geometry in updateCoords, evaluated in updateShifts.  It takes DFTB+'s native
!> arrays, coords(3,N) and latVecs(3,3) with the lattice vectors as columns, without transposes.
!>
!> Errors: the one-shot routines keep the reference's convention (`error stop` with a message,
!> fe/common/ewald_validation.f90:41-46); the engine routines return the C status instead.
module fftewald_amd
   use, intrinsic :: iso_c_binding, only: c_int, c_double, c_float, c_long_long, c_ptr, &
                                          c_null_ptr, c_char, c_associated, c_f_pointer, c_loc
   implicit none

   private
   public :: ew_options, ew_parameters, TEwaldEngine
   public :: ewald_potential_force, ewald_parameters, ewald_grid_memory
   public :: engine_create, engine_set_geometry, engine_shift, engine_destroy, engine_error
   public :: engine_add_gradients, engine_add_stress, engine_shift_multi, engine_get_potential
   public :: engine_add_external_pot_grad, engine_get_potential_gradient
   public :: ew_mixer_options, TEwaldMixer, mixer_create, mixer_reset, mixer_mix, mixer_destroy
   public :: EW_OK, EW_ERR_ARGUMENT, EW_PRECISION_FP64, EW_PRECISION_MIXED, EW_MIXER_BROYDEN
   public :: EW_MIXER_SIMPLE

   integer(c_int), parameter :: EW_OK = 0
   !> An array shorter than the resident geometry (the engine wrappers check the sizes that the
   !> C entry points take on trust), a null pointer, a bad option value.
   integer(c_int), parameter :: EW_ERR_ARGUMENT = 1
   integer(c_int), parameter :: EW_PRECISION_FP64 = 0
   integer(c_int), parameter :: EW_PRECISION_MIXED = 1
   integer(c_int), parameter :: EW_MIXER_BROYDEN = 0
   integer(c_int), parameter :: EW_MIXER_SIMPLE = 1

   !> Mirrors struct ew_tuning (launch shapes).
   type, bind(C) :: ew_tuning
      integer(c_int) :: spread_direct = 0
      integer(c_int) :: spread_block(3) = 0
      integer(c_int) :: fft_planes = 0
      integer(c_int) :: fft_lines(3) = 0
      integer(c_int) :: fft_radix(3) = 0
      integer(c_int) :: pair_cosched = 0
      integer(c_int) :: pair_spread_share = 0
      real(c_float) :: eval_us = 0.0_c_float
      integer(c_int) :: fft_wave = 0
      integer(c_int) :: fft_lanes(3) = 0
      integer(c_int) :: pair_blocks = 0
      integer(c_int) :: pair_fork = 0
   end type ew_tuning

   !> Mirrors struct ew_options.
   type, bind(C) :: ew_options
      real(c_double) :: tol = 0.0_c_double
      real(c_double) :: alpha = 0.0_c_double
      real(c_double) :: r_cut = 0.0_c_double
      integer(c_int) :: grid(3) = 0
      integer(c_int) :: order = 0
      integer(c_int) :: precision = 0
      integer(c_int) :: device = 0
      integer(c_int) :: alpha_rule = 0
      integer(c_int) :: profile = 0
      integer(c_int) :: eager = 0
      integer(c_int) :: charged = 0
      integer(c_int) :: fft_mode = 0
      integer(c_int) :: pair_storage = 0
      integer(c_int) :: group_grid = 0
      real(c_double) :: k_cut = 0.0_c_double
      integer(c_int) :: tune = 0
      type(ew_tuning) :: tuning = ew_tuning()
   end type ew_options

   !> Mirrors struct ew_parameters (the TEwaldParameters3d analogue).
   type, bind(C) :: ew_parameters
      real(c_double) :: alpha = 0.0_c_double
      real(c_double) :: r_cut = 0.0_c_double
      real(c_double) :: k_cut = 0.0_c_double
      real(c_double) :: r_list = 0.0_c_double
      integer(c_int) :: grid(3) = 0
      integer(c_int) :: order = 0
      integer(c_int) :: precision = 0
      integer(c_int) :: pair_mode = 0
      integer(c_int) :: index_bytes = 0
      integer(c_int) :: fft_planes = 0
      integer(c_int) :: gather_split = 0
      integer(c_long_long) :: n_pairs = 0
      integer(c_long_long) :: device_bytes = 0
   end type ew_parameters

   !> Mirrors struct ew_mixer_options (TBroydenMixerInp, broydenmixer.F90:31-50).
   type, bind(C) :: ew_mixer_options
      integer(c_int) :: kind = 0
      integer(c_int) :: max_iter = 0
      real(c_double) :: mix_param = 0.0_c_double
      real(c_double) :: omega0 = 0.0_c_double
      real(c_double) :: min_weight = 0.0_c_double
      real(c_double) :: max_weight = 0.0_c_double
      real(c_double) :: weight_fac = 0.0_c_double
      integer(c_int) :: device = 0
      integer(c_int) :: reserved(3) = 0
   end type ew_mixer_options

   !> Device charge mixer handle (the TMixerReal analogue).
   type :: TEwaldMixer
      type(c_ptr) :: handle = c_null_ptr
   end type TEwaldMixer

   !> Resident engine handle.
   type :: TEwaldEngine
      type(c_ptr) :: handle = c_null_ptr
      integer :: nAtom = 0
   end type TEwaldEngine

   interface
      integer(c_int) function c_ew_create(opts, h) bind(C, name="ew_create")
         import :: c_int, c_ptr, ew_options
         type(ew_options), intent(in) :: opts
         type(c_ptr), intent(out) :: h
      end function c_ew_create

      integer(c_int) function c_ew_destroy(h) bind(C, name="ew_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function c_ew_destroy

      type(c_ptr) function c_ew_last_error(h) bind(C, name="ew_last_error")
         import :: c_ptr
         type(c_ptr), value :: h
      end function c_ew_last_error

      integer(c_int) function c_ew_set_geometry(h, n, coords, latvecs, species, hubbard_u, &
                                                n_species, charges) bind(C, name="ew_set_geometry")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: n, n_species
         real(c_double), intent(in) :: coords(3, *), latvecs(3, 3)
         type(c_ptr), value :: species, hubbard_u, charges
      end function c_ew_set_geometry

      integer(c_int) function c_ew_shift(h, dq, shift, energy) bind(C, name="ew_shift")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         real(c_double), intent(in) :: dq(*)
         real(c_double), intent(out) :: shift(*)
         real(c_double), intent(out) :: energy
      end function c_ew_shift

      integer(c_int) function c_ew_gradients(h, dq, grad) bind(C, name="ew_gradients")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         real(c_double), intent(in) :: dq(*)
         real(c_double), intent(inout) :: grad(3, *)
      end function c_ew_gradients

      integer(c_int) function c_ew_stress(h, dq, stress) bind(C, name="ew_stress")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         real(c_double), intent(in) :: dq(*)
         real(c_double), intent(inout) :: stress(3, 3)
      end function c_ew_stress

      integer(c_int) function c_ew_external_potential(h, m, coords, charges, blur, eps, pot) &
         bind(C, name="ew_external_potential")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: m
         real(c_double), intent(in) :: coords(3, *)
         real(c_double), intent(in) :: charges(*)
         type(c_ptr), value :: blur
         real(c_double), value :: eps
         real(c_double), intent(out) :: pot(*)
      end function c_ew_external_potential

      integer(c_int) function c_ew_external_gradients(h, m, coords, charges, blur, dq, agrad, &
         egrad) bind(C, name="ew_external_gradients")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: m
         real(c_double), intent(in) :: coords(3, *)
         real(c_double), intent(in) :: charges(*)
         type(c_ptr), value :: blur
         type(c_ptr), value :: dq
         real(c_double), intent(inout) :: agrad(3, *)
         type(c_ptr), value :: egrad
      end function c_ew_external_gradients

      integer(c_int) function c_ew_potential_force(n, coords, charges, latvecs, opts, pot, force) &
         bind(C, name="ew_potential_force")
         import :: c_int, c_double, c_ptr, ew_options
         integer(c_int), value :: n
         real(c_double), intent(in) :: coords(3, *), charges(*), latvecs(3, 3)
         type(ew_options), intent(in) :: opts
         real(c_double), intent(out) :: pot(*)
         type(c_ptr), value :: force
      end function c_ew_potential_force

      integer(c_int) function c_ew_parameters_for(n, charges, latvecs, opts, out) &
         bind(C, name="ew_parameters_for")
         import :: c_int, c_double, ew_options, ew_parameters
         integer(c_int), value :: n
         real(c_double), intent(in) :: charges(*), latvecs(3, 3)
         type(ew_options), intent(in) :: opts
         type(ew_parameters), intent(out) :: out
      end function c_ew_parameters_for

      integer(c_int) function c_ew_grid_memory(n, charges, latvecs, opts, gigabytes) &
         bind(C, name="ew_grid_memory")
         import :: c_int, c_double, ew_options
         integer(c_int), value :: n
         real(c_double), intent(in) :: charges(*), latvecs(3, 3)
         type(ew_options), intent(in) :: opts
         real(c_double), intent(out) :: gigabytes
      end function c_ew_grid_memory

      integer(c_int) function c_ew_shift_multi(h, nvec, dq, shift, energy) &
            bind(C, name="ew_shift_multi")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: nvec
         real(c_double), intent(in) :: dq(*)
         real(c_double), intent(out) :: shift(*)
         real(c_double), intent(out) :: energy(*)
      end function c_ew_shift_multi

      integer(c_int) function c_ew_mixer_create(opts, m) bind(C, name="ew_mixer_create")
         import :: c_int, c_ptr, ew_mixer_options
         type(ew_mixer_options), intent(in) :: opts
         type(c_ptr), intent(out) :: m
      end function c_ew_mixer_create

      integer(c_int) function c_ew_mixer_destroy(m) bind(C, name="ew_mixer_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: m
      end function c_ew_mixer_destroy

      integer(c_int) function c_ew_mixer_reset(m, n) bind(C, name="ew_mixer_reset")
         import :: c_int, c_ptr
         type(c_ptr), value :: m
         integer(c_int), value :: n
      end function c_ew_mixer_reset

      integer(c_int) function c_ew_mixer_mix(m, q, qd) bind(C, name="ew_mixer_mix")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: m
         real(c_double), intent(inout) :: q(*)
         real(c_double), intent(in) :: qd(*)
      end function c_ew_mixer_mix

      integer(c_int) function c_strlen(s) bind(C, name="strlen")
         import :: c_int, c_ptr
         type(c_ptr), value :: s
      end function c_strlen
   end interface

contains

   !> Message of the last failure (of h, or of the last one-shot call when h is null).
   function engine_error(h) result(msg)
      type(c_ptr), intent(in) :: h
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: i, n

      p = c_ew_last_error(h)
      n = c_strlen(p)
      call c_f_pointer(p, chars, [n])
      allocate (character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function engine_error

   !> Options from the reference's optional overrides: nModes_in and windowCutoff_in become the
   !> SPME mesh and B-spline order, the method this engine runs in place of the NFFT; k_cut_in
   !> replaces the Fourier cutoff the mesh is derived from (ewald_fft_3d_parameters.f90:127-130).
   function options_from_overrides(tol, alpha_in, r_cut_in, k_cut_in, nModes_in, &
                                   windowCutoff_in) result(o)
      real(c_double), intent(in), optional :: tol, alpha_in, r_cut_in, k_cut_in
      integer, intent(in), optional :: nModes_in, windowCutoff_in
      type(ew_options) :: o

      o = ew_options()
      if (present(tol)) o%tol = tol
      if (present(alpha_in)) o%alpha = alpha_in
      if (present(r_cut_in)) o%r_cut = r_cut_in
      if (present(k_cut_in)) o%k_cut = k_cut_in
      if (present(nModes_in)) o%grid = nModes_in
      if (present(windowCutoff_in)) o%order = windowCutoff_in
   end function options_from_overrides

   !> fe/fft/ewald_fft_3d.f90:36-134 with the same arguments: positions(i,:) is atom i,
   !> latVecs(i,:) is a_i; pot(i) = phi(x_i); force(:,i) = -dU/dx_i when present.
   !> k_cut_in replaces the Fourier cutoff; the mesh follows from it unless nModes_in is given.
   subroutine ewald_potential_force(positions, charges, nParticle, latVecs, tol, pot, force, &
                                    alpha_in, r_cut_in, k_cut_in, nModes_in, windowCutoff_in)
      real(c_double), intent(in) :: positions(:, :)
      real(c_double), intent(in) :: charges(:)
      integer, intent(in) :: nParticle
      real(c_double), intent(in) :: latVecs(3, 3)
      real(c_double), intent(in), optional :: tol
      real(c_double), intent(out) :: pot(:)
      real(c_double), intent(out), optional, target :: force(:, :)
      real(c_double), intent(in), optional :: alpha_in, r_cut_in, k_cut_in
      integer, intent(in), optional :: nModes_in, windowCutoff_in

      type(ew_options) :: o
      real(c_double), allocatable :: coords(:, :)
      real(c_double) :: lat(3, 3)
      integer(c_int) :: st

      if (nParticle <= 0) error stop "fftewald_amd: nParticle must be positive"
      if (size(positions, 2) /= 3) error stop "fftewald_amd: positions must be (N,3)"
      if (size(positions, 1) < nParticle .or. size(charges) < nParticle) &
         error stop "fftewald_amd: too few positions or charges"
      if (size(pot) < nParticle) error stop "fftewald_amd: pot must be sized (N)"
      o = options_from_overrides(tol, alpha_in, r_cut_in, k_cut_in, nModes_in, windowCutoff_in)
      allocate (coords(3, nParticle))
      coords = transpose(positions(1:nParticle, :))
      lat = transpose(latVecs)
      if (present(force)) then
         if (size(force, 1) /= 3 .or. size(force, 2) < nParticle) &
            error stop "fftewald_amd: force must be sized (3,N)"
         st = c_ew_potential_force(int(nParticle, c_int), coords, charges, lat, o, pot, &
                                   c_loc(force))
      else
         st = c_ew_potential_force(int(nParticle, c_int), coords, charges, lat, o, pot, c_null_ptr)
      end if
      if (st /= EW_OK) then
         write (*, '(A,I0,2A)') "fftewald_amd: status ", st, ": ", engine_error(c_null_ptr)
         error stop "fftewald_amd: ewald_potential_force failed"
      end if
   end subroutine ewald_potential_force

   !> fe/fft/ewald_fft_3d.f90:142-191.  positions are not needed for the chain.
   subroutine ewald_parameters(positions, charges, nParticle, latVecs, tol, params, alpha_in, &
                               r_cut_in, k_cut_in, nModes_in, windowCutoff_in)
      real(c_double), intent(in) :: positions(:, :)
      real(c_double), intent(in) :: charges(:)
      integer, intent(in) :: nParticle
      real(c_double), intent(in) :: latVecs(3, 3)
      real(c_double), intent(in), optional :: tol
      type(ew_parameters), intent(out) :: params
      real(c_double), intent(in), optional :: alpha_in, r_cut_in, k_cut_in
      integer, intent(in), optional :: nModes_in, windowCutoff_in
      type(ew_options) :: o
      integer(c_int) :: st

      if (size(positions, 1) < nParticle) error stop "fftewald_amd: too few positions"
      o = options_from_overrides(tol, alpha_in, r_cut_in, k_cut_in, nModes_in, windowCutoff_in)
      st = c_ew_parameters_for(int(nParticle, c_int), charges, transpose(latVecs), o, params)
      if (st /= EW_OK) error stop "fftewald_amd: ewald_parameters failed"
   end subroutine ewald_parameters

   !> fe/fft/ewald_fft_3d.f90:197-244: predicted device memory in gigabytes.
   function ewald_grid_memory(charges, nParticle, latVecs, tol, alpha_in, r_cut_in, k_cut_in, &
                              nModes_in, windowCutoff_in) result(gigabytes)
      real(c_double), intent(in) :: charges(:)
      integer, intent(in) :: nParticle
      real(c_double), intent(in) :: latVecs(3, 3)
      real(c_double), intent(in), optional :: tol, alpha_in, r_cut_in, k_cut_in
      integer, intent(in), optional :: nModes_in, windowCutoff_in
      real(c_double) :: gigabytes
      type(ew_options) :: o
      integer(c_int) :: st

      o = options_from_overrides(tol, alpha_in, r_cut_in, k_cut_in, nModes_in, windowCutoff_in)
      st = c_ew_grid_memory(int(nParticle, c_int), charges, transpose(latVecs), o, gigabytes)
      if (st /= EW_OK) error stop "fftewald_amd: ewald_grid_memory failed"
   end function ewald_grid_memory

   !> Create a resident engine (TCoulomb_init, coulomb.F90:446-506).
   function engine_create(this, opts) result(st)
      type(TEwaldEngine), intent(out) :: this
      type(ew_options), intent(in) :: opts
      integer(c_int) :: st

      st = c_ew_create(opts, this%handle)
   end function engine_create

   !> Geometry step (updateCoords/updateLatVecs, coulomb.F90:509-620) with DFTB+'s native
   !> coords(3,N) and latVecs(3,3) (columns a_i).  species/hubbardU enable the short gamma:
   !> species(i) is DFTB+'s 1-based species index of atom i (species0 in TScc%updateCoords,
   !> scc.F90:368-400) and hubbardU(s) the U of species s; the C ABI takes 0-based indices, so
   !> the wrapper shifts a copy.  An index outside 1..size(hubbardU) or a species array shorter
   !> than the geometry returns EW_ERR_ARGUMENT.
   function engine_set_geometry(this, coords, latVecs, species, hubbardU, charges) result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: coords(:, :)
      real(c_double), intent(in) :: latVecs(3, 3)
      integer, intent(in), optional :: species(:)
      real(c_double), intent(in), optional, target :: hubbardU(:)
      real(c_double), intent(in), optional, target :: charges(:)
      integer(c_int) :: st
      type(c_ptr) :: ps, pu, pq
      integer(c_int) :: ns
      integer(c_int), allocatable, target :: species0(:)

      ps = c_null_ptr
      pu = c_null_ptr
      pq = c_null_ptr
      ns = 0
      if (present(species) .and. present(hubbardU) .and. size(hubbardU) > 0) then
         st = EW_ERR_ARGUMENT
         if (size(species) < size(coords, 2)) return
         if (any(species(1:size(coords, 2)) < 1) .or. &
             any(species(1:size(coords, 2)) > size(hubbardU))) return
         allocate (species0(size(coords, 2)))
         species0(:) = int(species(1:size(coords, 2)) - 1, c_int)
         ps = c_loc(species0)
         pu = c_loc(hubbardU)
         ns = int(size(hubbardU), c_int)
      end if
      if (present(charges)) pq = c_loc(charges)
      st = c_ew_set_geometry(this%handle, int(size(coords, 2), c_int), coords, latVecs, ps, pu, &
                             ns, pq)
      if (st == EW_OK) this%nAtom = size(coords, 2)
   end function engine_set_geometry

   !> One SCC iteration (updateShifts, coulomb.F90:786-880): shift = phi(dq), energy.
   !> TCoulomb%addGradients (coulomb.F90:705-708): gradient(3, N) += dU/dx of the Ewald energy
   function engine_add_gradients(this, deltaQAtom, gradient) result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: deltaQAtom(:)
      real(c_double), intent(inout) :: gradient(:, :)
      integer(c_int) :: st

      st = EW_ERR_ARGUMENT
      if (size(deltaQAtom) < this%nAtom .or. .not. atoms3(this, gradient)) return
      st = c_ew_gradients(this%handle, deltaQAtom, gradient)
   end function engine_add_gradients

   !> TCoulomb%addStress (coulomb.F90:721-757): stress(3, 3) += the Ewald stress (symmetric, so
   !> the row-major C layout and Fortran's column-major one agree)
   function engine_add_stress(this, deltaQAtom, stress) result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: deltaQAtom(:)
      real(c_double), intent(inout) :: stress(3, 3)
      integer(c_int) :: st

      st = EW_ERR_ARGUMENT
      if (size(deltaQAtom) < this%nAtom) return
      st = c_ew_stress(this%handle, deltaQAtom, stress)
   end function engine_add_stress

   !> TCoulomb%getPotential, periodic branch (coulomb.F90:3121-3160): potential(nAtom) of the
   !> point charges chargeCoords(3, m), charges(m) at the resident atoms, with the reference's
   !> optional Gaussian blurWidths(m) and real-space softening epsSoften.
   function engine_get_potential(this, chargeCoords, charges, potential, blurWidths, epsSoften) &
      result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: chargeCoords(:, :)
      real(c_double), intent(in) :: charges(:)
      real(c_double), intent(out) :: potential(:)
      real(c_double), intent(in), optional, target, contiguous :: blurWidths(:)
      real(c_double), intent(in), optional :: epsSoften
      integer(c_int) :: st
      type(c_ptr) :: pb
      real(c_double) :: eps

      st = EW_ERR_ARGUMENT
      if (.not. sources_ok(chargeCoords, charges, blurWidths) .or. size(potential) < this%nAtom) &
         return
      pb = c_null_ptr
      if (present(blurWidths)) pb = c_loc(blurWidths)
      eps = 0.0_c_double
      if (present(epsSoften)) eps = epsSoften
      st = c_ew_external_potential(this%handle, int(size(charges), c_int), chargeCoords, charges, &
         pb, eps, potential)
   end function engine_get_potential

   !> TCoulomb%addExternalPotGrad, periodic (coulomb.F90:3197-3235): atomGrads(3, nAtom) +=
   !> dq_i sum_s q_s grad ewald(x_i - y_s), extChargeGrads(3, m) -= the same per point charge.
   function engine_add_external_pot_grad(this, chargeCoords, charges, deltaQAtom, atomGrads, &
      extChargeGrads, blurWidths) result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: chargeCoords(:, :)
      real(c_double), intent(in) :: charges(:)
      real(c_double), intent(in), target, contiguous :: deltaQAtom(:)
      real(c_double), intent(inout) :: atomGrads(:, :)
      real(c_double), intent(inout), target, contiguous :: extChargeGrads(:, :)
      real(c_double), intent(in), optional, target, contiguous :: blurWidths(:)
      integer(c_int) :: st
      type(c_ptr) :: pb

      st = EW_ERR_ARGUMENT
      if (.not. sources_ok(chargeCoords, charges, blurWidths)) return
      if (size(deltaQAtom) < this%nAtom .or. .not. atoms3(this, atomGrads)) return
      if (size(extChargeGrads, 1) /= 3 .or. size(extChargeGrads, 2) < size(charges)) return
      pb = c_null_ptr
      if (present(blurWidths)) pb = c_loc(blurWidths)
      st = c_ew_external_gradients(this%handle, int(size(charges), c_int), chargeCoords, charges, &
         pb, c_loc(deltaQAtom), atomGrads, c_loc(extChargeGrads))
   end function engine_add_external_pot_grad

   !> TCoulomb%getPotentialGradient, periodic (coulomb.F90:3162-3195): gradient(3, nAtom) =
   !> sum_s q_s grad ewald(x_i - y_s) (overwritten, as the reference zeroes it first).
   function engine_get_potential_gradient(this, chargeCoords, charges, gradient, blurWidths) &
      result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: chargeCoords(:, :)
      real(c_double), intent(in) :: charges(:)
      real(c_double), intent(out) :: gradient(:, :)
      real(c_double), intent(in), optional, target, contiguous :: blurWidths(:)
      integer(c_int) :: st
      type(c_ptr) :: pb

      st = EW_ERR_ARGUMENT
      if (.not. sources_ok(chargeCoords, charges, blurWidths) .or. .not. atoms3(this, gradient)) &
         return
      pb = c_null_ptr
      if (present(blurWidths)) pb = c_loc(blurWidths)
      gradient(:, :) = 0.0_c_double
      st = c_ew_external_gradients(this%handle, int(size(charges), c_int), chargeCoords, charges, &
         pb, c_null_ptr, gradient, c_null_ptr)
   end function engine_get_potential_gradient

   function engine_shift(this, deltaQAtom, shift, energy) result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: deltaQAtom(:)
      real(c_double), intent(out) :: shift(:)
      real(c_double), intent(out) :: energy
      integer(c_int) :: st

      st = EW_ERR_ARGUMENT
      if (size(deltaQAtom) < this%nAtom .or. size(shift) < this%nAtom) return
      st = c_ew_shift(this%handle, deltaQAtom, shift, energy)
   end function engine_shift

   !> nVec SCC iterations on one geometry in one launch: deltaQ(nAtom, nVec) -> shift(nAtom,
   !> nVec), energy(nVec) (column-major (nAtom, nVec) is the C [nvec][n] layout).
   function engine_shift_multi(this, deltaQ, shift, energy) result(st)
      type(TEwaldEngine), intent(inout) :: this
      real(c_double), intent(in) :: deltaQ(:, :)
      real(c_double), intent(out) :: shift(:, :)
      real(c_double), intent(out) :: energy(:)
      integer(c_int) :: st

      ! the C layout [nvec][n] is column-major (nAtom, nVec) with exactly nAtom rows
      st = EW_ERR_ARGUMENT
      if (size(deltaQ, 1) /= this%nAtom .or. size(shift, 1) /= this%nAtom) return
      if (size(shift, 2) < size(deltaQ, 2) .or. size(energy) < size(deltaQ, 2)) return
      st = c_ew_shift_multi(this%handle, int(size(deltaQ, 2), c_int), deltaQ, shift, energy)
   end function engine_shift_multi

   !> TBroydenMixerReal_init / TSimpleMixerReal_init (broydenmixer.F90:114-140).
   function mixer_create(this, opts) result(st)
      type(TEwaldMixer), intent(out) :: this
      type(ew_mixer_options), intent(in) :: opts
      integer(c_int) :: st

      st = c_ew_mixer_create(opts, this%handle)
   end function mixer_create

   !> TMixer%reset (broydenmixer.F90:144-169).
   function mixer_reset(this, nElem) result(st)
      type(TEwaldMixer), intent(inout) :: this
      integer, intent(in) :: nElem
      integer(c_int) :: st

      st = c_ew_mixer_reset(this%handle, int(nElem, c_int))
   end function mixer_reset

   !> TMixer%mix (broydenmixer.F90:180-200): qInpResult in -> mixed charges out.
   function mixer_mix(this, qInpResult, qDiff) result(st)
      type(TEwaldMixer), intent(inout) :: this
      real(c_double), intent(inout) :: qInpResult(:)
      real(c_double), intent(in) :: qDiff(:)
      integer(c_int) :: st

      st = c_ew_mixer_mix(this%handle, qInpResult, qDiff)
   end function mixer_mix

   subroutine mixer_destroy(this)
      type(TEwaldMixer), intent(inout) :: this
      integer(c_int) :: st

      if (c_associated(this%handle)) st = c_ew_mixer_destroy(this%handle)
      this%handle = c_null_ptr
   end subroutine mixer_destroy

   !> gradient-like array of shape (3, >= nAtom)
   logical function atoms3(this, a)
      type(TEwaldEngine), intent(in) :: this
      real(c_double), intent(in) :: a(:, :)
      atoms3 = size(a, 1) == 3 .and. size(a, 2) >= this%nAtom
   end function atoms3

   !> point charges: coordinates (3, >= m), optional blur widths (>= m), m = size(charges)
   logical function sources_ok(chargeCoords, charges, blurWidths)
      real(c_double), intent(in) :: chargeCoords(:, :)
      real(c_double), intent(in) :: charges(:)
      real(c_double), intent(in), optional :: blurWidths(:)
      sources_ok = size(chargeCoords, 1) == 3 .and. size(chargeCoords, 2) >= size(charges)
      if (present(blurWidths)) sources_ok = sources_ok .and. size(blurWidths) >= size(charges)
   end function sources_ok

   subroutine engine_destroy(this)
      type(TEwaldEngine), intent(inout) :: this
      integer(c_int) :: st

      if (c_associated(this%handle)) st = c_ew_destroy(this%handle)
      this%handle = c_null_ptr
   end subroutine engine_destroy

end module fftewald_amd
