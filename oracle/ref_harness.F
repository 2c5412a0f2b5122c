c-----------------------------------------------------------------------
c     Test harness around the compiled reference (test infrastructure
c     only).  It stands in for the reference's driver program: it sets
c     the COMMON state that src/driver.f:38-52 sets for a single rank
c     (init_dim, init_mesh for an explicit nelx x nely x nelz box,
c     proxy_setupds, set_multiplicity, proxy_setup, h1mg_setup), then
c     calls the reference's own entry points and writes raw little-
c     endian binaries next to the output prefix.
c
c     stdin:  line 1  task   (setup | ax | dssum | cg | hsmg | bench)
c             line 2  nelx nely nelz nx1 niter ncalls
c             line 3  input file (raw doubles, n of them) or '-'
c             line 4  output prefix
c
c     Every glsc3 result is appended to $HK_GLSC3 (see ref_hooks.c).
c-----------------------------------------------------------------------
      program refharness
      include 'SIZE'
      include 'TOTAL'
      include 'SEMHAT'
      include 'mpif.h'

      common /mymask/cmask(-1:lx1*ly1*lz1*lelt)
      parameter (lxyz = lx1*ly1*lz1)
      parameter (lt=lxyz*lelt)

      common /hkbig/ x(lt),f(lt),r(lt),w(lt),p(lt),z(lt),c(lt)
     $             , g(6,lt), ur(lt),us(lt),ut(lt),wk(lt)
      real x,f,r,w,p,z,c,g,ur,us,ut,wk
      common /hkglo/ glo(lt)
      integer*8 glo

      character*16 task
      character*512 fin, fout
      integer e, n, niter, ncalls, icall, nxin
      real*8 t0, t1, hk_clock
      external hk_clock

      call iniproc(mpi_comm_world)

      read(5,*) task
      read(5,*) nelx, nely, nelz, nxin, niter, ncalls
      read(5,'(a)') fin
      read(5,'(a)') fout

c     init_dim (src/driver.f:103-115)
      nx1  = nxin
      ny1  = nx1
      nz1  = nx1
      ndim = ldim

c     init_mesh for np=1 and an explicit box (src/driver.f:117-194):
c     offs=0 and the element loop runs x fastest, so lglel(e)=e.
      nelt = nelx*nely*nelz
      if (nx1.gt.lx1 .or. nelt.gt.lelt .or. nelt.lt.1) then
         write(6,*) 'HK size error', nx1, lx1, nelt, lelt
         stop 2
      endif
      n = nx1*ny1*nz1*nelt
      cmask(-1) = 1.0
      do e=1,nelt
         lglel(e) = e
      enddo

      t0 = hk_clock()
      call proxy_setupds(gsh,nx1)
      call hkphase('gs_setup', t0)

c     set_multiplicity (src/driver.f:274-291)
      call rone(c,n)
      call gs_op(gsh,c,1,1,0)
      do i=1,n
         c(i) = 1./c(i)
      enddo

      call hkphase('multiplicity', t0)
      call proxy_setup(ah,bh,ch,dh,zh,wh,g)
      call hkphase('proxy_setup', t0)
      call h1mg_setup
      call hkphase('h1mg_setup', t0)

      flop_a  = 0
      flop_cg = 0

      if (task.eq.'setup') then
         call set_vert_box(glo,nx1)
         call hkwr(trim(fout)//'.dxm1', dxm1, nx1*nx1)
         call hkwr(trim(fout)//'.dxtm1', dxtm1, nx1*nx1)
         call hkwr(trim(fout)//'.zgm1', zgm1, nx1)
         call hkwr(trim(fout)//'.wxm1', wxm1, nx1)
         call hkwr(trim(fout)//'.ah', ah, nx1*nx1)
         call hkwr(trim(fout)//'.c', c, n)
         call hkwr(trim(fout)//'.g', g, 6*n)
         call hkwri8(trim(fout)//'.glo', glo, n)

      elseif (task.eq.'ax') then
c        Two applications of ax to the same input: the first one meets
c        cmask(-1)>=0 and builds the mask list (src/cg.f:230-253).
         call hkrd(fin, p, n)
         call ax(w,p,g,ur,us,ut,wk,n)
         call hkwr(trim(fout)//'.w1', w, n)
         call ax(w,p,g,ur,us,ut,wk,n)
         call hkwr(trim(fout)//'.w2', w, n)
         nm = int(cmask(0))
         call hkwr(trim(fout)//'.cmask', cmask(-1), nm+2)

      elseif (task.eq.'setf') then
c        The reference's own right-hand side, set_f (src/driver.f:88-101).
         call set_f(f,c,n)
         call hkwr(trim(fout)//'.f', f, n)

      elseif (task.eq.'dssum') then
         call hkrd(fin, p, n)
         call dssum(p)
         call hkwr(trim(fout)//'.u', p, n)

      elseif (task.eq.'cg' .or. task.eq.'bench') then
c        RHS as set_f does (src/driver.f:96-97): dssum, then col2 by c,
c        applied to the raw vector read from the input file.
         t0 = hk_clock()
         call hkrd(fin, f, n)
         call hkphase('read rhs', t0)
         call dssum(f)
         call col2 (f,c,n)
         if (task.eq.'cg') call hkwr(trim(fout)//'.f', f, n)
         call hkphase('rhs', t0)
         do icall=1,ncalls
            t0 = hk_clock()
            call cg(x,f,g,c,r,w,p,z,n,niter,flop_cg)
            t1 = hk_clock()
            write(6,9) icall, niter, n, t1-t0
    9       format('HK cg call ',i3,' niter ',i6,' n ',i12,
     $             ' seconds ',1pe22.14)
            if (task.eq.'cg') then
               write(fin,'(a,i1)') '.x', icall
               call hkwr(trim(fout)//trim(fin), x, n)
               write(fin,'(a,i1)') '.r', icall
               call hkwr(trim(fout)//trim(fin), r, n)
            endif
         enddo

      elseif (task.eq.'hsmg') then
c        One preconditioner application; h1mg_solve masks its rhs in
c        place, so both the output and the modified rhs are written.
         call hkrd(fin, r, n)
         call h1mg_solve(z,r,n)
         call hkwr(trim(fout)//'.z', z, n)
         call hkwr(trim(fout)//'.r', r, n)

      else
         write(6,*) 'HK unknown task ', task
         stop 3
      endif

      call hk_flush()
      write(6,*) 'HK done ', trim(task)
      end
c-----------------------------------------------------------------------
      subroutine hkphase(name,t0)
c     Progress line for a finished setup phase, flushed at once so that
c     a caller streaming the output sees it; t0 restarts the clock.
      character*(*) name
      real*8 t0, t1, hk_clock
      external hk_clock
      t1 = hk_clock()
      write(6,'(a,a,a,f10.3)') ' HK phase ', name, ' seconds ', t1-t0
      flush(6)
      t0 = t1
      return
      end
c-----------------------------------------------------------------------
      subroutine hkwr(name,a,n)
      character*(*) name
      real a(n)
      open(unit=21,file=name,access='stream',form='unformatted',
     $     status='replace')
      write(21) a
      close(21)
      return
      end
c-----------------------------------------------------------------------
      subroutine hkwri8(name,a,n)
      character*(*) name
      integer*8 a(n)
      open(unit=21,file=name,access='stream',form='unformatted',
     $     status='replace')
      write(21) a
      close(21)
      return
      end
c-----------------------------------------------------------------------
      subroutine hkrd(name,a,n)
      character*(*) name
      real a(n)
      open(unit=22,file=trim(name),access='stream',form='unformatted',
     $     status='old')
      read(22) a
      close(22)
      return
      end
c-----------------------------------------------------------------------
