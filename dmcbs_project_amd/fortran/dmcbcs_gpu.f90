! DMC driver on the MI355X engine: reads the DMCBS deck from standard input in the
! reference's positional order (dmcbcs.f90:43-66, then the module decks), generates or reads
! the initial walkers, runs the block/step loop of dmcbcs.f90:426-449 through the host
! modules step/stack/wavefunction/estimator, prints the block results as resstring does and
! writes the final walkers to fort.9 (dmcbcs.f90:548-567).
!
! Usage:  dmcbcs_gpu [fp64|mixed] [allparticle|singleparticle] [krefresh] < deck
! The defaults (fp64 allparticle) propagate exactly as the reference's step1.  Decks with
! idmc = 2 run Metropolis DMC, idmc = 0 variational Monte Carlo, and idmc = -1 adds the
! parameter optimisation (calder after every block, updatepar after each of optiter
! iterations).
program dmcbcs_gpu
   use engine
   use potential
   use wavefunction
   use stack
   use step
   use estimator
   use optimizer
   use operators
   use random
   use random2
   implicit none
   integer(kind=i8) :: irn,irn0
   integer(kind=i4) :: npart,nwalk,idmc,neq,nav,nstep,ndim,ntab,nbin(2)
   integer(kind=i4) :: i,j,k,it,nt,n1,idum,prec,movet,kref,optiter,iters
   real(kind=r8) :: ene
   logical :: vmc
   real(kind=r8) :: hbar,dt,etrial,etrial0,popup,popdn,dcut,dobal,wtcut,mass(2)
   real(kind=r8) :: el,mu,eunit,kunit,rho,pi,tau,tau0
   logical :: isite,units,fn,euclid,empty,operat
   character(len=90), allocatable :: answer(:)
   character(len=32) :: arg
   real(kind=r8) :: timebl0,timebl,rn(1),dummy
   real(kind=r8), allocatable :: xtemp(:)

   prec=DMC_FP64
   movet=DMC_ALLPARTICLE
   kref=0
   do i=1,command_argument_count()
      call get_command_argument(i,arg)
      select case (trim(arg))
         case ('fp64')
            prec=DMC_FP64
         case ('mixed')
            prec=DMC_MIXED
         case ('allparticle')
            movet=DMC_ALLPARTICLE
         case ('singleparticle')
            movet=DMC_SINGLEPARTICLE
         case default
            read (arg,*) kref
      end select
   enddo

! deck header (dmcbcs.f90:43-66)
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
   etrial0=etrial
   vmc=idmc.eq.0.or.idmc.eq.-1
   if (vmc) then
      prec=DMC_FP64
      movet=DMC_VMC
   endif
   if (idmc.eq.2) then
      prec=DMC_FP64
      movet=DMC_METROPOLIS
   endif
   call engine_deck(nwalk,etrial0,units,popup,popdn,prec,movet,kref,int(10*nwalk,i4))

! module set-up in the order of dmcbcs.f90:142-177
   el=0.0_r8
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
   etrial=etrial*eunit
   if (el.ne.0.0_r8) etrial=etrial*npart
! the four operator table sizes (operators.f90:33-41): g(r) and S(q) on the device
   call setoperators(ndim,nbin,el,operat,kunit)
   euclid=.false.
   optiter=1
   if (idmc.eq.-1) then
      call setupopt(npart,optiter)
      operat=.false.  ! no operators during the optimisation (dmcbcs.f90:173-176)
   endif
   call setstep(idmc,ndim,npart,nbin,hbar,dt,etrial,el,mass,wtcut,dcut,eunit,fn)
   if (vmc) then
! the variational estimators of dmcbcs.f90:180-186
      call setestnum(6)
      allocate(answer(6))
      call addest(1,'total energy')
      call addest(2,'kinetic energy')
      call addest(3,'potential energy')
      call addest(4,'external potential')
      call addest(5,'jackson-feenberg energy')
      call addest(6,'acceptance =')
   elseif (idmc.eq.2) then
! the Metropolis DMC estimators of dmcbcs.f90:204-218
      call setestnum(13)
      allocate(answer(13))
      call addest(1,'total energy')
      call addest(2,'kinetic energy')
      call addest(3,'potential energy')
      call addest(4,'node crossing')
      call addest(5,'acceptance')
      call addest(6,'|psit|/psig =')
      call addest(7,'psit =')
      call addest(8,'psig =')
      call addest(9,'sign(psit) =')
      call addest(10,'wt =')
      call addest(11,'wtg =')
      call addest(12,'sign(wtg) =')
      call addest(13,'sign(wtg)-weighted =')
   else
   call setestnum(14)
   allocate(answer(14))
   call addest(1,'total energy')
   call addest(2,'kinetic energy')
   call addest(3,'potential energy')
   call addest(4,'external potential')
   call addest(5,'node crossing')
   call addest(6,'growth energy')
   call addest(7,'|psit|/psig =')
   call addest(8,'psit =')
   call addest(9,'psig =')
   call addest(10,'sign(psit) =')
   call addest(11,'wt =')
   call addest(12,'wtg =')
   call addest(13,'sign(wtg) =')
   call addest(14,'sign(wtg)-weighted =')
   endif
   call create(2,int(10*nwalk,i4),npart,ndim,euclid)
   allocate(xtemp(ndim))

! initial walkers (dmcbcs.f90:237-301): the reference's generation from one random
! configuration, on the host side of the engine, or the walker file
   if (isite) then
      write (6,'(''Initial walkers from random configs'')')
! one configuration from the module stream, then per walker npart random swaps and the next
! ran2 state of the walker stream, as dmcbcs.f90:240-286 generates them
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
      call flush_stack
      tau0=0.0_r8
   else
      write (6,'(''Initial walkers from file'')')
      open (unit=9,file='fort.9',status='old')
      read (9,'(i10,f15.8)') n1,tau0
      write (6,'(''Number of walkers ='',t40,i10)') n1
      do k=1,n1
         read (9,'(6e15.7)') w1%x
         read (9,'(e15.7,2e25.15)') w1%weight,w1%psil0,w1%psigl0
         read (9,'(i20)') w1%irn
         call push(1,w1)
      enddo
      close (9)
      call flush_stack
   endif
   write (6,'(''Configurations at tau ='',t35,f15.10)') tau0
   tau=tau0
! the operators of the initial walkers (dmcbcs.f90:379-395)
   if (operat.and.tau.eq.0.0_r8.and.idmc.ne.-1.and.idmc.ne.0) then
      call addoperators
      call updateoperators
      call writeoperators(tau)
   endif

! optimisation iterations (dmcbcs.f90:234-539): after the first, the walkers' stored
! wavefunction values are recomputed with the new parameters (dmcbcs.f90:314-353)
   do iters=1,optiter
   if (iters.gt.1) then
      write (6,'(''optimization iteration number '',i10)') iters
      call reload_stack
   endif
! blocks of nstep steps (dmcbcs.f90:426-449)
   it=0
   do i=1,nav+neq
      if (i.eq.neq+1) then
         if (neq.gt.0) write (6,'(/,''equilibration done'')')
         call zerest
         if (operat) call zerooperators
         it=0
      endif
      it=it+1
      call cpu_time(timebl0)
      do j=1,nstep
         call step1(1,2,j.eq.nstep)
         if (idmc.gt.0) call branch(2,1,1.0_r8)
         nt=numstack(1)
         if (nt.gt.popup*nwalk.or.nt.lt.popdn*nwalk) call branch(1,2,1.0_r8*nwalk/nt)
         if (.not.vmc) tau=tau+dt
      enddo
      call cpu_time(timebl)
      nt=numstack(1)
      call update
      write (6,'(/,''iteration ='',t30,i14)') it
      write (6,'(''walker number = '',t30,i14)') nt
      answer=resstring(tau)
      write (6,'(a90)') (answer(k),k=1,size(answer))
      write (6,'(''Time for block ='',f10.3,'' seconds'')') timebl-timebl0
! operators of the population at the end of each block past equilibration
! (dmcbcs.f90:465-478)
      if (operat.and.i.gt.neq) then
         call addoperators
         call updateoperators
         call writeoperators(tau)
      endif
! dmcbcs.f90:492-503: calder over the walkers after every block past equilibration
      if (idmc.eq.-1.and.i.gt.neq) call calder
   enddo
   if (idmc.eq.-1) call updatepar(ene)
   enddo

! final walkers to fort.9 (dmcbcs.f90:548-567)
   if (nav+neq.ne.0) then
      nt=numstack(1)
      open (unit=9,file='fort.9',status='replace')
      write (9,'(i10,f15.8)') nt,tau
      do k=1,nt
         call pop(1,w1,empty)
         if (empty) exit
         write (9,'(6e15.7)') w1%x
         write (9,'(e15.7,2e25.15)') w1%weight,w1%psil,w1%psil
         write (9,'(i20,2e15.7)') w1%irn
      enddo
      close (9)
      call writewtisto
   endif
   call engine_stop
end program dmcbcs_gpu
