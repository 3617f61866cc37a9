! A fragment of a DMCBS driver written against the host modules the way dmcbcs.f90 writes it:
! the deck header (dmcbcs.f90:43-66) and module set-up, the initial walkers from one random
! configuration with the module random-number stream (dmcbcs.f90:237-301), the initial hpsi
! pass over the stack with the walker fields it reads and sets (dmcbcs.f90:314-345), calder
! per walker and updatepar (dmcbcs.f90:492-503,538) and the walker file (dmcbcs.f90:548-567).
!
! Usage:  walkers_fragment gen|all < deck
!   gen  walker generation and the walker file of the generated walkers only (no device
!        call): walkers.out (the fort.9 format) and gauss.out (gaussian(7) after setrn)
!   all  also the hpsi pass and the optimiser on the device: hpsi.out, params.out and fort.9
program walkers_fragment
   use engine
   use random
   use random2
   use potential
   use wavefunction
   use stack
   use step
   use optimizer
   implicit none
   integer(kind=i8) :: irn,irn0,s
   integer(kind=i4) :: npart,nwalk,idmc,neq,nav,nstep,ndim,ntab,nbin(2),i,j,k,n1,nt
   integer(kind=i4) :: istin,istout,idum,optiter,npar
   real(kind=r8) :: hbar,dt,etrial,popup,popdn,dcut,dobal,wtcut,mass(2),mu,el,rn(1),dummy
   real(kind=r8) :: ene,tau,g(7),par(32)
   real(kind=r8), allocatable :: xtemp(:)
   logical :: isite,units,fn,empty,all
   character(len=8) :: mode

   call get_command_argument(1,mode)
   all=trim(mode).eq.'all'
! dmcbcs.f90:43-66
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
   call engine_deck(nwalk,etrial,units,popup,popdn,DMC_FP64,DMC_VMC,0,int(10*nwalk,i4))
   el=0.0_r8
   call setpsi(ndim,nbin,el,mass,ntab)
   call setphi(ndim,ntab,nbin,el,mass,hbar/(2.0_r8*mu))
   do i=1,4
      read (5,*) idum   ! the operator table sizes (operators.f90:33-41)
   enddo
   if (idmc.eq.-1) call setupopt(npart,optiter)
   call setstep(idmc,ndim,npart,nbin,hbar,dt,etrial,el,mass,wtcut,dcut,1.0_r8,fn)
   call create(2,int(10*nwalk,i4),npart,ndim,.false.)
   allocate(xtemp(ndim))

! dmcbcs.f90:237-301 with isite
   irn=irn0
   call setrn(irn)
   w1%x=reshape(el*(randn(size(w1%x))-0.5_r8),shape(w1%x))
   open (unit=10,file='walkers.out',status='replace')
   write (10,'(i10,f15.8)') nwalk,0.0_r8
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
      write (10,'(6e15.7)') w1%x
      write (10,'(e15.7,2e25.15)') w1%weight,w1%psil0,w1%psigl0
      write (10,'(i20)') w1%irn
      if (all) call push(1,w1)
   enddo
   close (10)
   s=7_i8
   call setrn(s)
   g=gaussian(7)
   call savern(s)
   open (unit=10,file='gauss.out',status='replace')
   write (10,'(7es26.17)') g
   write (10,'(i20)') s
   close (10)
   if (.not.all) stop

! dmcbcs.f90:314-345: hpsi of every walker, the fields it sets, calder per walker
   istin=1
   istout=2
   n1=numstack(istin)
   open (unit=10,file='hpsi.out',status='replace')
   do i=1,n1
      call pop(istin,w1,empty)
      if (empty) exit
      call hpsi(w1)
      ene=(-hbar*w1%d2psi+w1%v)/eunit_of(el)/npart
      w1%wt0=w1%is
      w1%weight=exp(w1%psigl-w1%psil0)
      w1%weight=1.0_r8
      write (10,'(i20,5es26.17,2i4)') w1%irn,w1%psil,w1%d2psi,w1%v,w1%psigl,w1%d2psig, &
         w1%is,w1%isg
      write (10,'(3es26.17)') w1%dpsi
      write (10,'(3es26.17)') w1%dpsig
      if (idmc.eq.-1) call calder(w1,ene)
      call push(istout,w1)
   enddo
   close (10)
   istout=istin
   istin=3-istin
   if (idmc.eq.-1) then
      call updatepar(ene)
      call getdetparam(npar,par)
      open (unit=10,file='params.out',status='replace')
      write (10,'(es26.17)') ene
      write (10,'(i6)') npar
      write (10,'(es26.17)') par(1:npar)
      close (10)
   endif
! dmcbcs.f90:548-567
   nt=numstack(istin)
   tau=0.0_r8
   open (unit=9,file='fort.9',status='replace')
   write (9,'(i10,f15.8)') nt,tau
   do i=1,nt
      call pop(istin,w1,empty)
      if (.not.empty) then
         write (9,'(6e15.7)') w1%x
         write (9,'(e15.7,2e25.15)') w1%weight,w1%psil,w1%psigl
         write (9,'(i20,2e15.7)') w1%irn
      endif
   enddo
   close (9)
   call engine_stop
contains
! the energy unit of dmcbcs.f90:146-166 (units: the Fermi energy per particle of the box)
   function eunit_of(el) result(e)
   real(kind=r8) :: el,e,rho,kunit,pi
   pi=4.0_r8*atan(1.0_r8)
   rho=npart/el**ndim
   kunit=(ndim*pi**(ndim-1)*rho)**(1.0_r8/ndim)
   e=3.0_r8/(8.0_r8-ndim)*hbar/(2.0_r8*mu)*kunit**2
   end function eunit_of
end program walkers_fragment
