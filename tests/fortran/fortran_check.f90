!> Drives the engine from Fortran exactly as DFTB+ would: the reference-signature
!> ewald_potential_force (fe/fft/ewald_fft_3d.f90:36-76) and the resident engine with DFTB+'s
!> native coords(3,N) / latVecs(3,3).  Reads input.bin, writes output.bin (stream access).
!>   input : n, coords(3,n), charges(n), latvecs(3,3), tol, alpha, grid, order, nsp, species(n), U(nsp)
!>   output: pot(n), shift(n), energy
program fortran_check
   use, intrinsic :: iso_c_binding, only: c_double, c_int
   use fftewald_amd
   implicit none
   integer(c_int) :: n, grid, order, nsp, st
   real(c_double), allocatable :: coords(:, :), charges(:), pot(:), shift(:), hub(:)
   integer(c_int), allocatable :: species(:)
   real(c_double) :: lat(3, 3), tol, alpha, energy
   type(TEwaldEngine) :: engine
   type(ew_options) :: opts
   character(len=512) :: inpath, outpath
   integer :: u

   call get_command_argument(1, inpath)
   call get_command_argument(2, outpath)
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
   st = engine_set_geometry(engine, coords, lat, species, hub, charges)
   if (st /= EW_OK) then
      print *, engine_error(engine%handle)
      error stop "engine_set_geometry failed"
   end if
   st = engine_shift(engine, charges, shift, energy)
   if (st /= EW_OK) error stop "engine_shift failed"
   call engine_destroy(engine)

   open (newunit=u, file=trim(outpath), access="stream", form="unformatted", status="replace")
   write (u) pot, shift, energy
   close (u)
   print '(A,I0,A,ES22.14)', "fortran_check: n=", n, " energy=", energy
end program fortran_check
