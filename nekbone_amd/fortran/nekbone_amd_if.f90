! ISO_C_BINDING interfaces of the nekbone_amd library's Fortran-callable
! entry points (include/nekbone_amd.h).  The reference-named ones keep the
! reference's argument lists and by-reference convention (REAL = real*8,
! INTEGER = int32, glo_num = integer*8), so a host that uses implicit
! interfaces, as the reference's F77 code does, needs none of this; an
! F2003 host uses the module for checked calls.
module nekbone_amd_if
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function nekbone_init(device) bind(C, name="nekbone_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    ! binds the mesh the reference-named calls act on (stands in for the
    ! reference's COMMON blocks /dimn/, /dxyz/, /mymask/, /comm_handles/)
    subroutine nekbone_setup(nelt, nx1, nelx, nely, nelz, lglel, dxm1, gsh) &
        bind(C, name="nekbone_setup_")
      import :: c_int, c_double
      integer(c_int) :: nelt, nx1, nelx, nely, nelz, lglel(*), gsh
      real(c_double) :: dxm1(*)
    end subroutine
    subroutine nekbone_use_jacobi(on) bind(C, name="nekbone_use_jacobi_")
      import :: c_int
      integer(c_int) :: on
    end subroutine
    ! src/semhat.f:2-37
    subroutine semhat(a, b, c, d, z, w, n) bind(C, name="semhat_")
      import :: c_int, c_double
      real(c_double) :: a(*), b(*), c(*), d(*), z(*), w(*)
      integer(c_int) :: n
    end subroutine
    ! src/cg.f:2-84
    subroutine cg(x, f, g, c, r, w, p, z, n, niter, flop_cg) bind(C, name="cg_")
      import :: c_int, c_double
      real(c_double) :: x(*), f(*), g(*), c(*), r(*), w(*), p(*), z(*), flop_cg
      integer(c_int) :: n, niter
    end subroutine
    ! src/cg.f:86-94
    subroutine solvem(z, r, n) bind(C, name="solvem_")
      import :: c_int, c_double
      real(c_double) :: z(*), r(*)
      integer(c_int) :: n
    end subroutine
    ! src/cg.f:96-125
    subroutine ax(w, u, gxyz, ur, us, ut, wk, n) bind(C, name="ax_")
      import :: c_int, c_double
      real(c_double) :: w(*), u(*), gxyz(*), ur(*), us(*), ut(*), wk(*)
      integer(c_int) :: n
    end subroutine
    ! src/prox_dssum.f:2-12
    subroutine dssum(f) bind(C, name="dssum_")
      import :: c_double
      real(c_double) :: f(*)
    end subroutine
    ! src/math.f:603-617
    real(c_double) function glsc3(a, b, mult, n) bind(C, name="glsc3_")
      import :: c_int, c_double
      real(c_double) :: a(*), b(*), mult(*)
      integer(c_int) :: n
    end function
    ! src/hsmg.f:3-29 and :960-1034
    subroutine h1mg_setup() bind(C, name="h1mg_setup_")
    end subroutine
    subroutine h1mg_solve(z, rhs, n) bind(C, name="h1mg_solve_")
      import :: c_int, c_double
      real(c_double) :: z(*), rhs(*)
      integer(c_int) :: n
    end subroutine
    ! src/jl/gs.c:1160-1170, 1188-1193, 1237-1242
    subroutine gs_setup(h, glo_num, n, comm, np) bind(C, name="gs_setup_")
      import :: c_int, c_long_long
      integer(c_int) :: h, n, comm, np
      integer(c_long_long) :: glo_num(*)
    end subroutine
    subroutine gs_op(h, u, dom, op, transpose) bind(C, name="gs_op_")
      import :: c_int, c_double
      integer(c_int) :: h, dom, op, transpose
      real(c_double) :: u(*)
    end subroutine
    subroutine gs_free(h) bind(C, name="gs_free_")
      import :: c_int
      integer(c_int) :: h
    end subroutine
  end interface
end module
