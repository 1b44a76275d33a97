! C-callable entry points into the reference's own powerI4.f90, compiled unchanged from
! /root/reference/pbi4/F90_modules/powerI4.f90 by oracle/Makefile into oracle/_ref/.
! TEST INFRASTRUCTURE ONLY: this shim lets the CPU tests check the oracle restatement against
! the reference routines (assign_single_grid, assign_double_grid, fcomb, reduce_array, compute_q2k,
! compute_q4k, measure_pk, measure_pk_multipoles).  The product library never loads it.
! Arrays are passed as plain pointers with explicit sizes; the reference routines take
! assumed-shape dummies, so each call goes through an explicit interface.

module ref_bind
  use iso_c_binding
  implicit none

  interface
    subroutine assign_single_grid(grid, nintp, npart, pos, weight, dtl)
      integer, intent(in) :: grid, nintp, npart
      real(kind=8), intent(in) :: pos(:,:), weight(:)
      real(kind=8), intent(out) :: dtl(grid, grid, grid)
    end subroutine
    subroutine assign_double_grid(grid, nintp, npart, pos, weight, dcl)
      integer, intent(in) :: grid, nintp, npart
      real(kind=8), intent(in) :: pos(:,:), weight(:)
      complex(kind=8), intent(out) :: dcl(grid, grid, grid)
    end subroutine
    subroutine fcomb(grid, nintp, interlacing, box, fdcl, dcl)
      integer, intent(in) :: grid, nintp
      real, intent(in) :: box
      logical, intent(in) :: interlacing
      complex(kind=8), intent(in) :: fdcl(:,:,:)
      complex(kind=8), intent(out) :: dcl(grid, grid, grid)
    end subroutine
    subroutine reduce_array(L, din, grid, dout)
      integer, intent(in) :: L, grid
      complex(kind=8), intent(in) :: din(:,:,:)
      complex(kind=8), intent(out) :: dout(grid, grid, grid)
    end subroutine
    subroutine compute_q2k(grid, box, deltak, i, j, q2k)
      integer, intent(in) :: grid, i, j
      real, intent(in) :: box
      complex(kind=8), intent(in) :: deltak(:,:,:)
      complex(kind=8), intent(out) :: q2k(grid, grid, grid)
    end subroutine
    subroutine compute_q4k(grid, box, deltak, i, j, l, k, q4k)
      integer, intent(in) :: grid, i, j, l, k
      real, intent(in) :: box
      complex(kind=8), intent(in) :: deltak(:,:,:)
      complex(kind=8), intent(out) :: q4k(grid, grid, grid)
    end subroutine
    subroutine measure_pk(grid, box, kbin, kcenter, nbins, dcl1, dcl2, pout)
      integer, intent(in) :: grid, nbins
      real, intent(in) :: box, kbin, kcenter
      complex(kind=8), intent(in) :: dcl1(:,:,:), dcl2(:,:,:)
      real(kind=8), intent(out) :: pout(5, nbins)
    end subroutine
    subroutine measure_pk_multipoles(grid, box, kbin, kcenter, nbins, dcl, dcl2, dcl4, pout)
      integer, intent(in) :: grid, nbins
      real, intent(in) :: box, kbin, kcenter
      complex(kind=8), intent(in) :: dcl(:,:,:), dcl2(:,:,:), dcl4(:,:,:)
      real(kind=8), intent(out) :: pout(5, nbins)
    end subroutine
  end interface

contains

  subroutine ref_assign_single_grid(grid, nintp, npart, pos, weight, dtl) bind(C)
    integer(c_int), value :: grid, nintp, npart
    real(c_double), intent(in) :: pos(3, npart), weight(npart)
    real(c_double), intent(out) :: dtl(grid, grid, grid)
    dtl = 0.d0
    call assign_single_grid(grid, nintp, npart, pos, weight, dtl)
  end subroutine

  subroutine ref_assign_double_grid(grid, nintp, npart, pos, weight, dcl) bind(C)
    integer(c_int), value :: grid, nintp, npart
    real(c_double), intent(in) :: pos(3, npart), weight(npart)
    complex(c_double_complex), intent(out) :: dcl(grid, grid, grid)
    dcl = (0.d0, 0.d0)
    call assign_double_grid(grid, nintp, npart, pos, weight, dcl)
  end subroutine

  subroutine ref_fcomb(grid, nintp, interlacing, box, fdcl, dcl) bind(C)
    integer(c_int), value :: grid, nintp, interlacing
    real(c_float), value :: box
    complex(c_double_complex), intent(in) :: fdcl(grid, grid, grid)
    complex(c_double_complex), intent(out) :: dcl(grid, grid, grid)
    logical :: il
    il = interlacing /= 0
    call fcomb(grid, nintp, il, box, fdcl, dcl)
  end subroutine

  subroutine ref_reduce_array(L, din, grid, dout) bind(C)
    integer(c_int), value :: L, grid
    complex(c_double_complex), intent(in) :: din(L, L, L)
    complex(c_double_complex), intent(out) :: dout(grid, grid, grid)
    dout = (0.d0, 0.d0)
    call reduce_array(L, din, grid, dout)
  end subroutine

  subroutine ref_compute_q2k(grid, box, deltak, i, j, q2k) bind(C)
    integer(c_int), value :: grid, i, j
    real(c_float), value :: box
    complex(c_double_complex), intent(in) :: deltak(grid, grid, grid)
    complex(c_double_complex), intent(out) :: q2k(grid, grid, grid)
    call compute_q2k(grid, box, deltak, i, j, q2k)
  end subroutine

  subroutine ref_compute_q4k(grid, box, deltak, i, j, l, k, q4k) bind(C)
    integer(c_int), value :: grid, i, j, l, k
    real(c_float), value :: box
    complex(c_double_complex), intent(in) :: deltak(grid, grid, grid)
    complex(c_double_complex), intent(out) :: q4k(grid, grid, grid)
    call compute_q4k(grid, box, deltak, i, j, l, k, q4k)
  end subroutine

  subroutine ref_measure_pk(grid, box, kbin, kcenter, nbins, dcl1, dcl2, pout) bind(C)
    integer(c_int), value :: grid, nbins
    real(c_float), value :: box, kbin, kcenter
    complex(c_double_complex), intent(in) :: dcl1(grid, grid, grid), dcl2(grid, grid, grid)
    real(c_double), intent(out) :: pout(5, nbins)
    call measure_pk(grid, box, kbin, kcenter, nbins, dcl1, dcl2, pout)
  end subroutine

  subroutine ref_measure_pk_multipoles(grid, box, kbin, kcenter, nbins, dcl, dcl2, dcl4, pout) bind(C)
    integer(c_int), value :: grid, nbins
    real(c_float), value :: box, kbin, kcenter
    complex(c_double_complex), intent(in) :: dcl(grid, grid, grid), dcl2(grid, grid, grid)
    complex(c_double_complex), intent(in) :: dcl4(grid, grid, grid)
    real(c_double), intent(out) :: pout(5, nbins)
    call measure_pk_multipoles(grid, box, kbin, kcenter, nbins, dcl, dcl2, dcl4, pout)
  end subroutine

end module ref_bind
