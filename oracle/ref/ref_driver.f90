! Test-infrastructure driver for the reference DMCBS modules.
!
! This program is part of the oracle: it is linked against the reference's own
! hot-path modules (ran.f90, kshell.f90, rungekutta.f90, betafun.f90,
! coshpot.f90, jastrowlongrange.f90, psislaterbcs.f90, matrixlapack.f90,
! stack.f90, estimator.f90, wavefunction.f90, step.f90, optimizer.f90 and the
! reference's single-process nompi.f90), compiled from the reference tree by
! oracle/Makefile into oracle/_ref/.  It is never part of the product.
!
! The reference program dmcbcs.f90 cannot be linked against nompi.f90 (it calls
! checkpop/zerotimebal/printmpilog, which nompi.f90 lacks), so this driver
! restates the parts of dmcbcs.f90 it needs:
!   * deck header read            dmcbcs.f90:43-66
!   * module set-up order         dmcbcs.f90:142-177
!   * energy unit                 dmcbcs.f90:146-166
!   * initial walkers (isite)     dmcbcs.f90:237-301
!   * initial hpsi pass           dmcbcs.f90:314-353
!   * main DMC step loop          dmcbcs.f90:426-449
! and then calls the reference's own setpsi/setphi/setstep/hpsi/step1/branch/
! move1/getderpsi/update/result unchanged.
!
! The observables modes also link the reference's gofr.f90, sofq.f90 and momentum3d.f90.
!
! Usage:  ref_driver <mode> <infile> <outfile>  < deck
! All binary files are little-endian stream files of int64 / float64 values.
program ref_driver
   use stack
   use random
   use random2
   use wavefunction
   use estimator
   use step
   use mympi
   use potential
   use betafun
   use kshell
   use optimizer
   use gofr
   use sofq
   use pdistribution
   implicit none
   integer, parameter :: i4=selected_int_kind(9)
   integer, parameter :: i8=selected_int_kind(15)
   integer, parameter :: r8=selected_real_kind(15,9)
   integer(kind=i8) :: irn,irn0
   integer(kind=i4) :: npart,nwalk,idmc,neq,nav,nstep,ndim,ntab,nbin(2)
   real(kind=r8) :: hbar,dt,etrial,popup,popdn,dcut,dobal,wtcut,mass(2)
   logical :: isite,units,fn,operat,euclid
   real(kind=r8) :: el,mu,eunit,kunit,rho,pi,nfac
   character(len=32) :: mode
   character(len=512) :: infile,outfile
   integer(kind=i4) :: idum,optiter
   el=0.0_r8
   call init0
   call get_command_argument(1,mode)
   call get_command_argument(2,infile)
   call get_command_argument(3,outfile)
! deck header, same order as dmcbcs.f90:43-66
   read (5,*) irn
   irn0=irn
   read (5,*) idmc
   read (5,*) isite
   read (5,*) nwalk
   read (5,*) neq
   read (5,*) nav
   read (5,*) nstep
   read (5,*) dt
   read (5,*) etrial
   read (5,*) units
   read (5,*) dcut
   read (5,*) fn
   read (5,*) popup
   read (5,*) popdn
   read (5,*) dobal
   read (5,*) wtcut
   read (5,*) ntab
   read (5,*) hbar
   read (5,*) ndim
   read (5,*) npart
   read (5,*) nbin(1),mass(1)
   read (5,*) nbin(2),mass(2)
   mu=mass(1)*mass(2)/(mass(1)+mass(2))
! module set-up, same order as dmcbcs.f90:142-177
   call setpsi(ndim,nbin,el,mass,ntab)
   call setphi(ndim,ntab,nbin,el,mass,hbar/(2.0_r8*mu))
   if (el.ne.0.0_r8.and.units) then
      rho=npart/el**(ndim)
      pi=4.0_r8*atan(1.0_r8)
      kunit=(ndim*pi**(ndim-1)*rho)**(1.0_r8/ndim)
      eunit=3.0_r8/(8.0_r8-ndim)*hbar/(2.0_r8*mu)*kunit**2
   else
      eunit=1.0_r8
      kunit=1.0_r8
   endif
   nfac=1.0_r8
   etrial=etrial*eunit
   if (el.ne.0.0_r8) then
      etrial=etrial*npart
      nfac=1.0_r8/npart
   endif
! operator switches (setoperators reads four lines); operators are not used
   read (5,*) idum
   read (5,*) idum
   read (5,*) idum
   read (5,*) idum
   operat=.false.
   euclid=.false.
   if (idmc.eq.-1.or.idmc.eq.-2) then
      optiter=1
      call setupopt(npart,optiter)
   endif
   call setstep(idmc,ndim,npart,nbin,hbar,dt,etrial,el,mass,wtcut,dcut,eunit,fn)
   call setestnum(14)
   call create(2,int(nwalk*10.0_r8),npart,ndim,euclid)

   select case (trim(mode))
      case ('setup')
         call do_setup
      case ('hpsi')
         call do_hpsi
      case ('ratio')
         call do_ratio
      case ('der')
         call do_der
      case ('dmc')
         call do_dmc(.false.)
      case ('dmcfile')
         call do_dmc(.true.)
      case ('opt')
         call do_opt
      case ('obs')
         call do_obs
      case ('enk')
         call do_enk
      case default
         write (6,*) 'unknown mode ',trim(mode)
         stop 1
   end select

contains

   subroutine openout
   open (unit=21,file=trim(outfile),access='stream',form='unformatted',status='replace')
   end subroutine openout

   subroutine openin
   open (unit=22,file=trim(infile),access='stream',form='unformatted',status='old')
   end subroutine openin

! set-up quantities: box, energy unit, k-vectors, radial tables at given r,
! closed-form beta, derbeta, vpot; RNG streams.
   subroutine do_setup
   integer(kind=i8) :: nr,i,nseed,ns,j,ng
   real(kind=r8), allocatable :: rr(:)
   real(kind=r8) :: f,df,d2f,b1,b1p,b1pp,dbet(2),rn,g(7)
   integer(kind=i4) :: norb,nk,numsh(64)
   real(kind=r8), pointer :: ak(:,:),ak2(:)
   integer(kind=i8) :: s1,s2,seeds(16),norb8
   call openin
   read (22) norb8
   norb=int(norb8,i4)
   read (22) nr
   allocate(rr(nr))
   read (22) rr
   read (22) nseed
   read (22) seeds(1:nseed)
   read (22) ns
   close (22)
   call openout
   write (21) el,eunit,etrial
   call setupk(el,norb,numsh,ak,ak2,nk,ndim)
   write (21) int(nk,i8)
   write (21) int(numsh(1:norb),i8)
   write (21) ak(:,1:nk)
   write (21) ak2(1:nk)
   do i=1,nr
      call radialfun(rr(i),f,df,d2f)
      call beta(rr(i),b1,b1p,b1pp)
      call derbeta(rr(i),dbet)
      write (21) f,df,d2f,b1,b1p,b1pp,dbet(1),dbet(2),vpot(rr(i))
   enddo
! raw LCG streams: ran1 and ran2 from each seed
   do j=1,nseed
      s1=seeds(j)
      s2=seeds(j)
      do i=1,ns
         call ran1(rn,s1)
         write (21) s1,rn
         call ran2(rn,s2)
         write (21) s2,rn
      enddo
   enddo
! gaussian(7) (odd length uses the cos branch) from each seed
   do j=1,nseed
      call setrn(seeds(j))
      g=gaussian(7)
      write (21) g
      call savern(s1)
      write (21) s1
   enddo
   close (21)
   end subroutine do_setup

   subroutine read_config(x)
   real(kind=r8) :: x(:,:)
   read (22) x
   end subroutine read_config

! getphi, getpsi and hpsi on configurations read from infile
   subroutine do_hpsi
   integer(kind=i8) :: nc,k
   real(kind=r8) :: u,d2u,v,phil,d2phi
   real(kind=r8), allocatable :: du(:,:),dphi(:,:)
   integer(kind=i4) :: isj,isd
   allocate(du(ndim,npart),dphi(ndim,npart))
   call openin
   read (22) nc
   call openout
   do k=1,nc
      call read_config(w1%x)
      call getphi(w1%x,u,isj,du,d2u,v)
      call getpsi(w1%x,phil,isd,dphi,d2phi)
      call hpsi(w1)
      write (21) u,d2u,v,phil,d2phi,real(isd,r8)
      write (21) du,dphi
      write (21) w1%psil,w1%d2psi,w1%v,real(w1%is,r8),w1%psigl,w1%d2psig
      write (21) w1%dpsi,w1%dpsig
   enddo
   close (21)
   close (22)
   end subroutine do_hpsi

! single-particle ratios psi(x')/psi(x) through move1 (getphione*bcsone)
   subroutine do_ratio
   integer(kind=i8) :: nc,k,nm,m,ip
   real(kind=r8) :: xnew(3),psirat
   call openin
   read (22) nc
   read (22) nm
   call openout
   do k=1,nc
      call read_config(w1%x)
      call hpsi(w1)
      do m=1,nm
         read (22) ip
         read (22) xnew
         call move1(int(ip,i4),xnew,w1,psirat)
         write (21) psirat
      enddo
   enddo
   close (21)
   close (22)
   end subroutine do_ratio

! d(log psi)/d(alpha_s) through getderpsi (getderbcs)
   subroutine do_der
   integer(kind=i8) :: nc,k
   integer(kind=i4) :: npar
   real(kind=r8), pointer :: par(:)
   real(kind=r8), allocatable :: dp(:)
   call getdetparam(npar,par)
   allocate(dp(npar))
   call openin
   read (22) nc
   call openout
   write (21) int(npar,i8)
   write (21) par
   do k=1,nc
      call read_config(w1%x)
      call getderpsi(w1%x,dp)
      write (21) dp
   enddo
   close (21)
   close (22)
   end subroutine do_der

! VMC optimiser accumulation: hpsi + calder on configurations, then updatepar
   subroutine do_opt
   integer(kind=i8) :: nc,k
   real(kind=r8) :: ene
   integer(kind=i4) :: npar
   real(kind=r8), pointer :: par(:)
   call openin
   read (22) nc
   do k=1,nc
      call read_config(w1%x)
      call hpsi(w1)
      ene=(-hbar*w1%d2psi+w1%v)/eunit/npart
      call calder(w1,ene)
   enddo
   close (22)
   call updatepar(ene)
! the optimiser's new parameters into the wavefunction (optimizer.f90:73-88), as the next
! optimisation iteration of dmcbcs.f90:236 does
   call updateparams
   call getdetparam(npar,par)
   call openout
   write (21) ene
   write (21) int(npar,i8)
   write (21) par
   close (21)
   end subroutine do_opt

! write every walker of stack istin (order as popped) to unit 21
   subroutine dump_stack(istin,istout)
   integer(kind=i4) :: istin,istout,nw,i
   logical :: empty
   nw=numstack(istin)
   write (21) int(nw,i8)
   do i=1,nw
      call pop(istin,w1,empty)
      write (21) w1%irn
      write (21) w1%weight,w1%psil,w1%d2psi,w1%v,real(w1%is,r8),real(w1%wt0,r8)
      write (21) w1%x
      write (21) w1%dpsi
      call push(istout,w1)
   enddo
   istout=istin
   istin=3-istin
   end subroutine dump_stack

! full DMC run: initial walkers, initial hpsi, then nav*nstep steps of
! step1 + branch + population control as in dmcbcs.f90:426-449.
! infile holds the list of steps after which the walker stack is dumped
! (and, for mode dmcfile, the initial walkers).
   subroutine do_dmc(fromfile)
   logical :: fromfile
   integer(kind=i4) :: istin,istout,n1,nt,i,j,k,jj,ist
   integer(kind=i8) :: ndump,nin,kk
   integer(kind=i8), allocatable :: dumps(:)
   real(kind=r8) :: rn(1),dummy,tau,valn,errn,val,err,ene,timebl0,timebl
   real(kind=r8), allocatable :: xtemp(:)
   real(kind=r8), pointer :: par(:)
   integer(kind=i4) :: npar
   logical :: empty
   allocate(xtemp(ndim))
   call openin
   read (22) ndump
   allocate(dumps(ndump))
   read (22) dumps
   call openout
   if (fromfile) then
      read (22) nin
      do kk=1,nin
         read (22) w1%x
         read (22) w1%irn
         w1%weight=1.0_r8
         w1%x0=0.0_r8
         call push(1,w1)
      enddo
   else
! dmcbcs.f90:237-301 (isite=.true., one process)
      irn=irn0
      call setrn(irn)
      w1%x=reshape(el*(randn(size(w1%x))-0.5_r8),shape(w1%x))
      do k=1,nwalk
         do i=1,npart
            rn=randn(1)
            j=npart*rn(1)+1
            xtemp(:)=w1%x(:,i)
            w1%x(:,i)=w1%x(:,j)
            w1%x(:,j)=xtemp(:)
         enddo
         w1%weight=1.0_r8
         w1%irn=irn
         call ran2(dummy,w1%irn)
         irn=w1%irn
         w1%x0=0.0_r8
         call push(1,w1)
      enddo
   endif
   close (22)
   istin=1
   istout=2
   n1=numstack(istin)
! dmcbcs.f90:314-353
   do i=1,n1
      call pop(istin,w1,empty)
      if (empty) exit
      call hpsi(w1)
      w1%wt0=w1%is
      w1%weight=1.0_r8
      call push(istout,w1)
   enddo
   istout=istin
   istin=3-istin
   call zerest
   tau=0.0_r8
   jj=0
   write (21) int(nav*nstep,i8)
   if (any(dumps.eq.0)) call dump_stack(istin,istout)
   do i=1,nav
! the block timer of dmcbcs.f90:434,450,487 (process CPU time around the block's steps),
! printed in the reference's format
      call cpu_time(timebl0)
      do j=1,nstep
         jj=jj+1
         call step1(istin,istout,j.eq.nstep)
! dmcbcs.f90:441: no branching in variational runs
         if (idmc.gt.0.or.idmc.eq.-2) call branch(istin,istout,1.0_r8)
         nt=numstack(istin)
         write (21) int(nt,i8)
         if (nt.gt.popup*nwalk.or.nt.lt.popdn*nwalk) then
            call branch(istin,istout,1.0_r8*nwalk/nt)
         endif
         nt=numstack(istin)
         write (21) int(nt,i8)
         tau=tau+dt
         if (any(dumps.eq.jj)) call dump_stack(istin,istout)
      enddo
      call cpu_time(timebl)
      write (6,'(''Time for block (sec) (min) = '',t35,2f11.3)') timebl-timebl0, &
         (timebl-timebl0)/60.0_r8
      call update
      do k=1,14
         call result(k,valn,errn,val,err)
         write (21) valn,errn
      enddo
! dmcbcs.f90:492-503: optimisation runs accumulate calder over the walkers after each block
      if (idmc.eq.-1.or.idmc.eq.-2) then
         do while(.true.)
            call pop(istin,w1,empty)
            if (empty) exit
            call hpsi(w1)
            ene=(-hbar*w1%d2psi+w1%v)/eunit/npart
            call calder(w1,ene)
            call push(istout,w1)
         enddo
         istout=istin
         istin=3-istout
      endif
   enddo
! dmcbcs.f90:538: the parameter update at the end of the optimisation iteration
   if (idmc.eq.-1.or.idmc.eq.-2) then
      call updatepar(ene)
      call updateparams
      call getdetparam(npar,par)
      write (21) ene
      write (21) int(npar,i8)
      write (21) par
   endif
   close (21)
   end subroutine do_dmc

! g(r), the density profile and S(q) of configurations read with their weights from infile:
! the reference's setupgofr/addgofr/updategofr/writegofr (gofr.f90:14-176) and
! setupsofq/addsofq/updatesofq/writesofq (sofq.f90:14-118) unchanged, as operators.f90:66-67,
! 103-104, 120-121 and 137-138 call them; the tables go to gofr.out, rho.out and sofq.out in
! the working directory, outfile gets the number of configurations, el and k_F.
   subroutine do_obs
   integer(kind=i8) :: nc,k,ntg,ntq
   real(kind=r8) :: wt
   call openin
   read (22) ntg
   read (22) ntq
   read (22) nc
   call setupgofr(ndim,int(ntg,i4),nbin,el,kunit)
   call setupsofq(ndim,nbin,el,int(ntq,i4),kunit)
   do k=1,nc
      call read_config(w1%x)
      read (22) wt
      call addgofr(w1%x,wt)
      call addsofq(w1%x,wt)
   enddo
   close (22)
   call updategofr
   call updatesofq
   call writegofr('gofr.out',0.0_r8)
   call writesofq('sofq.out',0.0_r8)
   call openout
   write (21) real(nc,r8),el,kunit
   close (21)
   end subroutine do_obs

! n(k) of configurations read with their RNG states and weights from infile: the
! reference's setupenk/zeroenk/addenk/updateenk/writeenk (momentum3d.f90:13-190) unchanged,
! as operators.f90:68,86,105,122,139 call them; the table goes to nofk.out in the working
! directory, outfile gets the number of configurations, el and k_F.
   subroutine do_enk
   integer(kind=i8) :: nc,k,ntk
   real(kind=r8) :: wt
   call openin
   read (22) ntk
   read (22) nc
   call setupenk(ndim,int(ntk,i4),el,kunit)
   call zeroenk
   do k=1,nc
      call read_config(w1%x)
      read (22) w1%irn
      read (22) wt
      call addenk(w1,wt)
   enddo
   close (22)
   call updateenk
   call writeenk('nofk.out',0.0_r8)
   call openout
   write (21) real(nc,r8),el,kunit
   close (21)
   end subroutine do_enk

end program ref_driver
