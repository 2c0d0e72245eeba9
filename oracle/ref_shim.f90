!> C-callable entry points onto the reference's own Fortran modules (direct Ewald,
!> alpha search, truncation, Lambert W, geometry, linked-cell real space), compiled
!> from the reference sources where they lie.  TEST INFRASTRUCTURE ONLY: the
!> tests use it to pin the C restatement in oracle/ewald_oracle.c.
!>
!> C layouts: positions [n][3] (Fortran (3,n)), lattice [3][3] with row i = a_i
!> (Fortran (3,3) with column i = a_i).  The reference wants positions(N,3) and
!> latVecs(3,3) with row i = a_i, so both are transposed here, as the reference's
!> own adapter does (src/dftbp/dftb/coulomb.F90:343-347).
module ref_shim
   use iso_c_binding, only: c_int, c_double
   use ewald_constants, only: dp, pi
   use ewald_geometry, only: TCell3d, cell_metrics_3d
   use ewald_direct_3d, only: direct_potential_force => ewald_potential_force, &
                              balanced_splitting
   use ewald_alpha_search, only: balanced_alpha
   use ewald_truncation, only: direct_cutoffs, splitting_from_real_budget, &
                               cutoff_from_fourier_budget
   use ewald_lambert_w, only: lambert_w0
   use ewald_fft_3d_real, only: short_range_potential_force
   implicit none

contains

   subroutine ref_direct_ewald(n, pos_c, q, lat_c, tol, alpha_in, want_force, pot, force) &
      bind(C, name="ref_direct_ewald")
      integer(c_int), value :: n, want_force
      real(c_double), intent(in) :: pos_c(3, n), q(n), lat_c(3, 3)
      real(c_double), value :: tol, alpha_in
      real(c_double), intent(out) :: pot(n), force(3, n)
      real(dp), allocatable :: positions(:, :)
      real(dp) :: latVecs(3, 3)

      allocate (positions(n, 3))
      positions = transpose(pos_c)
      latVecs = transpose(lat_c)
      if (alpha_in > 0.0_dp) then
         if (want_force /= 0) then
            call direct_potential_force(positions, q, n, latVecs, tol, pot, force, alpha_in=alpha_in)
         else
            call direct_potential_force(positions, q, n, latVecs, tol, pot, alpha_in=alpha_in)
         end if
      else
         if (want_force /= 0) then
            call direct_potential_force(positions, q, n, latVecs, tol, pot, force)
         else
            call direct_potential_force(positions, q, n, latVecs, tol, pot)
         end if
      end if
   end subroutine ref_direct_ewald

   subroutine ref_short_range(n, pos_c, q, lat_c, alpha, r_cut, pot, force) &
      bind(C, name="ref_short_range")
      integer(c_int), value :: n
      real(c_double), intent(in) :: pos_c(3, n), q(n), lat_c(3, 3)
      real(c_double), value :: alpha, r_cut
      real(c_double), intent(out) :: pot(n), force(3, n)
      real(dp), allocatable :: positions(:, :)
      type(TCell3d) :: cell

      allocate (positions(n, 3))
      positions = transpose(pos_c)
      cell = cell_metrics_3d(transpose(lat_c))
      call short_range_potential_force(positions, q, n, cell, alpha, r_cut, pot, force)
   end subroutine ref_short_range

   subroutine ref_cell_metrics(lat_c, rec_c, volume, latlen, reclen) &
      bind(C, name="ref_cell_metrics")
      real(c_double), intent(in) :: lat_c(3, 3)
      real(c_double), intent(out) :: rec_c(3, 3), volume, latlen(3), reclen(3)
      type(TCell3d) :: cell

      cell = cell_metrics_3d(transpose(lat_c))
      rec_c = transpose(cell%recVecs)
      volume = cell%volume
      latlen = cell%latLengths
      reclen = cell%recLengths
   end subroutine ref_cell_metrics

   subroutine ref_balanced_alpha(minG, minR, volume, tol, alpha, status) &
      bind(C, name="ref_balanced_alpha")
      real(c_double), value :: minG, minR, volume, tol
      real(c_double), intent(out) :: alpha
      integer(c_int), intent(out) :: status
      integer :: st

      call balanced_alpha(minG, minR, volume, tol, alpha, st)
      status = st
   end subroutine ref_balanced_alpha

   function ref_balanced_splitting(lat_c, tol) result(alpha) bind(C, name="ref_balanced_splitting")
      real(c_double), intent(in) :: lat_c(3, 3)
      real(c_double), value :: tol
      real(c_double) :: alpha

      alpha = balanced_splitting(cell_metrics_3d(transpose(lat_c)), tol)
   end function ref_balanced_splitting

   subroutine ref_direct_cutoffs(alpha, tol, r_cut, k_cut) bind(C, name="ref_direct_cutoffs")
      real(c_double), value :: alpha, tol
      real(c_double), intent(out) :: r_cut, k_cut

      call direct_cutoffs(alpha, tol, r_cut, k_cut)
   end subroutine ref_direct_cutoffs

   function ref_lambert_w0(x) result(w) bind(C, name="ref_lambert_w0")
      real(c_double), value :: x
      real(c_double) :: w

      w = lambert_w0(x)
   end function ref_lambert_w0

   function ref_splitting_from_real_budget(r_cut, q2, n, pv, budget) result(alpha) &
      bind(C, name="ref_splitting_from_real_budget")
      real(c_double), value :: r_cut, q2, pv, budget
      integer(c_int), value :: n
      real(c_double) :: alpha

      alpha = splitting_from_real_budget(r_cut, q2, int(n), pv, budget)
   end function ref_splitting_from_real_budget

   function ref_cutoff_from_fourier_budget(q2, n, edge, alpha, budget) result(k_cut) &
      bind(C, name="ref_cutoff_from_fourier_budget")
      real(c_double), value :: q2, edge, alpha, budget
      integer(c_int), value :: n
      real(c_double) :: k_cut

      k_cut = cutoff_from_fourier_budget(q2, int(n), edge, alpha, budget)
   end function ref_cutoff_from_fourier_budget

end module ref_shim
