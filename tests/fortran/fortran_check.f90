!> Drives the engine from Fortran exactly as DFTB+ would: the reference-signature
!> ewald_potential_force (fe/fft/ewald_fft_3d.f90:36-76) and the resident engine with DFTB+'s
!> native coords(3,N) / latVecs(3,3).  Reads input.bin, writes output.bin (stream access).
!>   input : n, coords(3,n), charges(n), latvecs(3,3), tol, alpha, grid, order, nsp, species(n), U(nsp)
!>   output: pot(n), shift(n), energy
!> With a third argument "mixer" it drives the device Broyden mixer as DFTB+'s SCC loop drives
!> TMixer%mix instead:
!>   input : nElem, nIter, maxIter, mixParam, omega0, minWeight, maxWeight, weightFac,
!>           q0(nElem), qDiff(nElem, nIter)
!>   output: qInpResult(nElem, nIter) after each mix
!> With "params" it calls the reference-signature ewald_parameters without and with k_cut_in
!> (no device needed):
!>   input : n, charges(n), latvecs(3,3), tol, kcut
!>   output: grid(3) without, grid(3) with k_cut_in, k_cut(2) reported
!> With "external" it drives the external-charge calls as TCoulomb%getPotential,
!> %getPotentialGradient and %addExternalPotGrad would be driven:
!>   input : n, coords(3,n), charges(n), latvecs(3,3), tol, m, chargeCoords(3,m), q(m), blur(m), eps
!>   output: pot(n), potBlurSoft(n), atomGrads(3,n), extGrads(3,m), potGrad(3,n)
program fortran_check
   use, intrinsic :: iso_c_binding, only: c_double, c_int
   use fftewald_amd
   implicit none
   integer(c_int) :: n, grid, order, nsp, st
   real(c_double), allocatable :: coords(:, :), charges(:), pot(:), shift(:), hub(:)
   integer, allocatable :: species(:)
   real(c_double) :: lat(3, 3), tol, alpha, energy
   type(TEwaldEngine) :: engine
   type(ew_options) :: opts
   character(len=512) :: inpath, outpath, mode
   integer :: u

   call get_command_argument(1, inpath)
   call get_command_argument(2, outpath)
   if (command_argument_count() >= 3) then
      call get_command_argument(3, mode)
      if (trim(mode) == "external") then
         call external_check(trim(inpath), trim(outpath))
      else if (trim(mode) == "params") then
         call params_check(trim(inpath), trim(outpath))
      else
         call mixer_check(trim(inpath), trim(outpath))
      end if
      stop
   end if
   open (newunit=u, file=trim(inpath), access="stream", form="unformatted", status="old")
   read (u) n
   allocate (coords(3, n), charges(n), pot(n), shift(n), species(n))
   read (u) coords, charges, lat, tol, alpha, grid, order, nsp
   allocate (hub(nsp))
   read (u) species, hub
   close (u)

   ! the reference's own signature: positions(N,3), lattice vectors as rows
   call ewald_potential_force(transpose(coords), charges, n, transpose(lat), tol, pot, &
                              alpha_in=alpha, nModes_in=grid, windowCutoff_in=order)

   ! the resident engine, as TCoulomb/TShortGamma would hold it
   opts%tol = tol
   opts%alpha = alpha
   opts%grid = grid
   opts%order = order
   st = engine_create(engine, opts)
   if (st /= EW_OK) error stop "engine_create failed"
   ! input.bin holds 0-based species; DFTB+ numbers them from 1, and so does the wrapper, which
   ! hands the C ABI 0-based indices.  An index out of range is refused, not read as another
   ! species' U.
   if (nsp > 0 .and. any(species == 0)) then
      if (engine_set_geometry(engine, coords, lat, species, hub, charges) /= EW_ERR_ARGUMENT) &
         error stop "engine_set_geometry accepted species index 0 (1-based input)"
   end if
   st = engine_set_geometry(engine, coords, lat, species + 1, hub, charges)
   if (st /= EW_OK) then
      print *, engine_error(engine%handle)
      error stop "engine_set_geometry failed"
   end if
   ! arrays shorter than the geometry are refused before the C side reads them
   if (engine_shift(engine, charges(1:size(charges) - 1), shift, energy) /= EW_ERR_ARGUMENT) &
      error stop "engine_shift accepted a short charge array"
   st = engine_shift(engine, charges, shift, energy)
   if (st /= EW_OK) error stop "engine_shift failed"
   call engine_destroy(engine)

   open (newunit=u, file=trim(outpath), access="stream", form="unformatted", status="replace")
   write (u) pot, shift, energy
   close (u)
   print '(A,I0,A,ES22.14)', "fortran_check: n=", n, " energy=", energy

contains

   subroutine params_check(inpath, outpath)
      character(len=*), intent(in) :: inpath, outpath
      integer(c_int) :: n
      real(c_double), allocatable :: charges(:), positions(:, :)
      real(c_double) :: lat(3, 3), tol, kcut
      type(ew_parameters) :: p0, p1
      integer :: u

      open (newunit=u, file=inpath, access="stream", form="unformatted", status="old")
      read (u) n
      allocate (charges(n), positions(n, 3))
      read (u) charges, lat, tol, kcut
      close (u)
      positions = 0.0_c_double
      call ewald_parameters(positions, charges, n, transpose(lat), tol, p0)
      call ewald_parameters(positions, charges, n, transpose(lat), tol, p1, k_cut_in=kcut)
      open (newunit=u, file=outpath, access="stream", form="unformatted", status="replace")
      write (u) p0%grid, p1%grid, p0%k_cut, p1%k_cut
      close (u)
      print '(A,3I5,A,3I5)', "fortran_check params: grid", p0%grid, " with k_cut_in", p1%grid
   end subroutine params_check

   subroutine mixer_check(inpath, outpath)
      character(len=*), intent(in) :: inpath, outpath
      integer(c_int) :: nElem, nIter, maxIter, st
      real(c_double) :: prm(5)
      real(c_double), allocatable :: q(:), qDiff(:, :), hist(:, :)
      type(ew_mixer_options) :: mopts
      type(TEwaldMixer) :: mixer
      integer :: u, it

      open (newunit=u, file=inpath, access="stream", form="unformatted", status="old")
      read (u) nElem, nIter, maxIter, prm
      allocate (q(nElem), qDiff(nElem, nIter), hist(nElem, nIter))
      read (u) q, qDiff
      close (u)
      mopts%kind = EW_MIXER_BROYDEN
      mopts%max_iter = maxIter
      mopts%mix_param = prm(1)
      mopts%omega0 = prm(2)
      mopts%min_weight = prm(3)
      mopts%max_weight = prm(4)
      mopts%weight_fac = prm(5)
      st = mixer_create(mixer, mopts)
      if (st /= EW_OK) error stop "mixer_create failed"
      st = mixer_reset(mixer, nElem)
      if (st /= EW_OK) error stop "mixer_reset failed"
      do it = 1, nIter
         st = mixer_mix(mixer, q, qDiff(:, it))
         if (st /= EW_OK) error stop "mixer_mix failed"
         hist(:, it) = q
      end do
      call mixer_destroy(mixer)
      open (newunit=u, file=outpath, access="stream", form="unformatted", status="replace")
      write (u) hist
      close (u)
      print '(A,I0,A,I0)', "fortran_check mixer: nElem=", nElem, " nIter=", nIter
   end subroutine mixer_check

   subroutine external_check(inpath, outpath)
      character(len=*), intent(in) :: inpath, outpath
      integer(c_int) :: n, m, st
      real(c_double), allocatable :: coords(:, :), charges(:), ext(:, :), qext(:), blur(:)
      real(c_double), allocatable :: pot(:), potb(:), agrad(:, :), egrad(:, :), pgrad(:, :)
      real(c_double) :: lat(3, 3), tol, eps
      type(TEwaldEngine) :: engine
      type(ew_options) :: opts
      integer :: u

      open (newunit=u, file=inpath, access="stream", form="unformatted", status="old")
      read (u) n
      allocate (coords(3, n), charges(n), pot(n), potb(n), agrad(3, n), pgrad(3, n))
      read (u) coords, charges, lat, tol, m
      allocate (ext(3, m), qext(m), blur(m), egrad(3, m))
      read (u) ext, qext, blur, eps
      close (u)
      opts%tol = tol
      st = engine_create(engine, opts)
      if (st /= EW_OK) error stop "engine_create failed"
      st = engine_set_geometry(engine, coords, lat, charges=charges)
      if (st /= EW_OK) error stop "engine_set_geometry failed"
      st = engine_get_potential(engine, ext, qext, pot)
      if (st /= EW_OK) error stop "engine_get_potential failed"
      st = engine_get_potential(engine, ext, qext, potb, blurWidths=blur, epsSoften=eps)
      if (st /= EW_OK) error stop "engine_get_potential (blur) failed"
      agrad(:, :) = 0.0_c_double
      egrad(:, :) = 0.0_c_double
      st = engine_add_external_pot_grad(engine, ext, qext, charges, agrad, egrad)
      if (st /= EW_OK) error stop "engine_add_external_pot_grad failed"
      st = engine_get_potential_gradient(engine, ext, qext, pgrad)
      if (st /= EW_OK) error stop "engine_get_potential_gradient failed"
      call engine_destroy(engine)
      open (newunit=u, file=outpath, access="stream", form="unformatted", status="replace")
      write (u) pot, potb, agrad, egrad, pgrad
      close (u)
      print '(A,I0,A,I0)', "fortran_check external: n=", n, " m=", m
   end subroutine external_check

end program fortran_check
