! A Fortran 2003 host of the nekbone_amd library.  It follows the
! reference driver's call sequence for one mesh (src/driver.f:38-62) and
! reaches the GPU only through the reference-named entry points:
!   init_mesh      lglel(e) = e (one rank, src/driver.f:154-191)
!   proxy_setupds  global numbering (set_vert_box, src/prox_dssum.f:42-73)
!                  and gs_setup
!   set_multiplicity  c = 1 / dssum(1) (src/driver.f:274-291)
!   proxy_setup    semhat -> dxm1, GLL weights; brick g (src/prox_setup.f)
!   then nekbone_setup, which binds the mesh for cg/ax/dssum/glsc3
! and runs tasks read from stdin.
!
! stdin:  line 1   nelx nely nelz nx1 niter
!         then one task per line, "<task> <input> <output prefix>", ending
!         with "end".  <input> holds n raw little-endian doubles.
!   cg     on a freshly bound mesh, f = c * dssum(input) (set_f's
!          post-processing, src/driver.f:96-97); cg twice as the driver
!          does (src/driver.f:54,59), the second timed; writes <prefix>.x
!          and <prefix>.r
!   ax     on a freshly bound mesh (the reference's mask state after
!          init_mesh), ax twice on the input (the first builds the
!          reference's mask list); writes <prefix>.w1 and <prefix>.w2
!   dssum  dssum of the input; writes <prefix>.u
!   glsc3  prints glsc3(input, c, input)
!   hsmg   h1mg_setup, then h1mg_solve(z, input); writes <prefix>.z and
!          the input as h1mg_solve leaves it, <prefix>.r
program nekbone_host
  use iso_c_binding
  use nekbone_amd_if
  implicit none
  integer(c_int) :: nelx, nely, nelz, nx1, niter, nelt, n, nxyz, gsh, comm, np, nm1
  integer(c_int) :: e, i, j, k, ios
  integer(c_int), allocatable :: lglel(:)
  integer(c_long_long), allocatable :: glo(:)
  real(c_double), allocatable :: x(:), f(:), g(:), c(:), r(:), w(:), p(:), z(:), u(:)
  real(c_double), allocatable :: ur(:), us(:), ut(:), wk(:)
  real(c_double) :: ah(256), bh(16), ch(256), dh(256), zh(16), wh(32), flop_cg, wt, s
  character(len=16) :: task
  character(len=512) :: fin, fout
  character(len=1600) :: line
  integer(8) :: t0, t1, rate

  read(*, *) nelx, nely, nelz, nx1, niter
  nelt = nelx * nely * nelz
  nxyz = nx1 * nx1 * nx1
  n = nxyz * nelt
  allocate(lglel(nelt), glo(n), x(n), f(n), g(6 * n), c(n), r(n), w(n), p(n), z(n), u(n))
  allocate(ur(n), us(n), ut(n), wk(n))

  do e = 1, nelt
    lglel(e) = e
  end do
  call set_vert_box(glo, nx1, nelx, nely, nelt)
  comm = 0
  np = 1
  call gs_setup(gsh, glo, n, comm, np)

  nm1 = nx1 - 1
  call semhat(ah, bh, ch, dh, zh, wh, nm1)
  call bind_mesh()
  call set_brick_g(g, bh, nx1, nelt)
  c = 1.0d0
  call dssum(c)
  c = 1.0d0 / c

  do
    ! whole line, split at blanks (list-directed input would end the
    ! record at the first '/' of a path)
    read(*, '(a)', iostat=ios) line
    if (ios /= 0) exit
    call next_word(line, task)
    call next_word(line, fin)
    call next_word(line, fout)
    if (task == 'end') exit
    select case (trim(task))
    case ('cg')
      call bind_mesh()
      call read_raw(fin, f, n)
      call dssum(f)
      f = f * c
      flop_cg = 0.0d0
      call cg(x, f, g, c, r, w, p, z, n, niter, flop_cg)
      call system_clock(t0, rate)
      call cg(x, f, g, c, r, w, p, z, n, niter, flop_cg)
      call system_clock(t1)
      wt = dble(t1 - t0) / dble(rate)
      write(*, '(a,i12,a,i6,a,1pe14.6,a,1pe22.14)') 'nekbone_host: n ', n, ' niter ', niter, &
          ' flop_cg ', flop_cg, ' seconds ', wt
      call write_raw(trim(fout) // '.x', x, n)
      call write_raw(trim(fout) // '.r', r, n)
    case ('ax')
      call bind_mesh()
      call read_raw(fin, u, n)
      call ax(w, u, g, ur, us, ut, wk, n)
      call write_raw(trim(fout) // '.w1', w, n)
      call ax(w, u, g, ur, us, ut, wk, n)
      call write_raw(trim(fout) // '.w2', w, n)
    case ('dssum')
      call read_raw(fin, u, n)
      call dssum(u)
      call write_raw(trim(fout) // '.u', u, n)
    case ('glsc3')
      call read_raw(fin, u, n)
      s = glsc3(u, c, u, n)
      write(*, '(a,1pe24.16)') 'nekbone_host: glsc3 ', s
    case ('hsmg')
      call h1mg_setup()
      call read_raw(fin, u, n)
      call h1mg_solve(z, u, n)
      call write_raw(trim(fout) // '.z', z, n)
      call write_raw(trim(fout) // '.r', u, n)
    case default
      write(*, *) 'nekbone_host: unknown task ', trim(task)
      stop 3
    end select
  end do
  call gs_free(gsh)
  write(*, '(a)') 'nekbone_host: done'

contains

  ! First blank-separated word of s into w; s loses it.
  subroutine next_word(s, w)
    character(len=*), intent(inout) :: s
    character(len=*), intent(out) :: w
    integer :: i0, i1
    s = adjustl(s)
    i1 = index(s, ' ')
    if (i1 == 0) i1 = len(s) + 1
    i0 = 1
    w = s(i0:i1 - 1)
    s = s(i1:)
  end subroutine

  subroutine bind_mesh()
    call nekbone_setup(nelt, nx1, nelx, nely, nelz, lglel, dh, gsh)
  end subroutine

  ! Global node numbers of the box (src/prox_dssum.f:42-73), one rank.
  subroutine set_vert_box(glo_num, nx, nelx, nely, nelt)
    integer(c_long_long), intent(out) :: glo_num(*)
    integer(c_int), intent(in) :: nx, nelx, nely, nelt
    integer(c_long_long) :: ii, kg, jg, ig
    integer(c_int) :: e, i, j, k, ex, ey, ez, nn, nelxy, ll
    nn = nx - 1
    nelxy = nelx * nely
    do e = 1, nelt
      ez = 1 + (e - 1) / nelxy
      ey = 1 + mod(e - 1, nelxy) / nelx
      ex = 1 + mod(e - 1, nelx)
      do k = 0, nn
        do j = 0, nn
          do i = 0, nn
            kg = int(nn * (ez - 1) + k, c_long_long)
            jg = int(nn * (ey - 1) + j, c_long_long)
            ig = int(nn * (ex - 1) + i, c_long_long)
            ii = 1 + ig + jg * (nn * nelx + 1) + kg * (nn * nelx + 1) * (nn * nely + 1)
            ll = 1 + i + nx * j + nx * nx * k + nx * nx * nx * (e - 1)
            glo_num(ll) = ii
          end do
        end do
      end do
    end do
  end subroutine

  ! Brick geometric factors (src/prox_setup.f:27-52): g1 = g4 = g6 =
  ! w_i w_j w_k, the others zero; g(6, nx1^3 nelt).
  subroutine set_brick_g(g, wx, nx, nelt)
    real(c_double), intent(out) :: g(6, *)
    real(c_double), intent(in) :: wx(*)
    integer(c_int), intent(in) :: nx, nelt
    integer(c_int) :: e, i, j, k, q
    do e = 1, nelt
      do k = 1, nx
        do j = 1, nx
          do i = 1, nx
            q = i + nx * (j - 1) + nx * nx * (k - 1) + nx * nx * nx * (e - 1)
            g(:, q) = 0.0d0
            g(1, q) = wx(i) * wx(j) * wx(k)
            g(4, q) = wx(i) * wx(j) * wx(k)
            g(6, q) = wx(i) * wx(j) * wx(k)
          end do
        end do
      end do
    end do
  end subroutine

  subroutine read_raw(name, a, n)
    character(len=*), intent(in) :: name
    integer(c_int), intent(in) :: n
    real(c_double), intent(out) :: a(n)
    open(unit=22, file=trim(name), access='stream', form='unformatted', status='old')
    read(22) a
    close(22)
  end subroutine

  subroutine write_raw(name, a, n)
    character(len=*), intent(in) :: name
    integer(c_int), intent(in) :: n
    real(c_double), intent(in) :: a(n)
    open(unit=21, file=trim(name), access='stream', form='unformatted', status='replace')
    write(21) a
    close(21)
  end subroutine

end program
