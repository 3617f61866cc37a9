! Fortran host modules over the MI355X engine.  They keep the module interfaces of the
! reference's walker-propagation path so that a driver written against them reads like
! dmcbcs.f90:
!   potential    setphi                          (coshpot.f90:14-143)
!   wavefunction setpsi, hpsi, getderpsi, setdetparam, getdetparam, move1
!                                                (wavefunction.f90:13-404)
!   random2      ran1, ran2                                  (ran.f90:1-50)
!   random       randn, gaussian, setrn, savern              (ran.f90:52-124)
!   stack        type(walker), create, push, pop, numstack   (stack.f90:6-119)
!   step         setstep, step1, branch, writewtisto         (step.f90:11-334)
!   estimator    setestnum, addest, zerest, update, resstring (estimator.f90:19-149)
!   optimizer    setupopt, calder(w, ene0) and calder over the population, updatepar
!                                                (optimizer.f90:14-169)
!   gofr, sofq, pdistribution   setup, zero, add, update, write of g(r), S(q), n(k)
!                (gofr.f90:14-177, sofq.f90:14-119, momentum3d.f90:13-190)
!   operators    setoperators, zero/add/update/writeoperators (operators.f90:7-140, g(r),
!                S(q) and n(k))
! All computation runs on the device through the C-ABI (module dmc_gpu); the walker
! population stays in device memory.  The deck values each reference module reads from
! unit 5 are read here in the same order and collected into one dmc_params record, and the
! device context is created on first use after setstep.
module engine
   use iso_c_binding
   use dmc_gpu
   implicit none
   integer, parameter :: i4=selected_int_kind(9)
   integer, parameter :: i8=selected_int_kind(15)
   integer, parameter :: r8=selected_real_kind(15,9)
   type(dmc_params), save :: prm
   type(c_ptr), save :: ctx=c_null_ptr
   logical, save :: started=.false.
contains
! run-level deck values that the reference keeps in dmcbcs.f90 (raw etrial in units of the
! Fermi energy per particle when units is set, as read at dmcbcs.f90:51-52)
   subroutine engine_deck(nwalk,etrial,units,popup,popdn,precision,movetype,krefresh,capacity)
   integer(kind=i4) :: nwalk,precision,movetype,krefresh,capacity
   real(kind=r8) :: etrial,popup,popdn
   logical :: units
   prm%nwalk=nwalk
   prm%etrial=etrial
   prm%units=merge(1,0,units)
   prm%popup=popup
   prm%popdn=popdn
   prm%precision=precision
   prm%movetype=movetype
   prm%krefresh=krefresh
   prm%capacity=capacity
   end subroutine engine_deck

   function engine_ctx() result(c)
   type(c_ptr) :: c
   if (.not.started) then
      call dmc_check(dmc_create(prm,ctx),'dmc_create')
      started=.true.
   endif
   c=ctx
   end function engine_ctx

   subroutine engine_stop
   integer(c_int) :: ierr
   if (started) ierr=dmc_destroy(ctx)
   started=.false.
   ctx=c_null_ptr
   end subroutine engine_stop
end module engine

module potential
   use engine
   implicit none
contains
! setphi: reads v0, amu, cjas(1:2), djas, v0op, v0eq (coshpot.f90:35-41); the tables are
! built by the engine from these values
   subroutine setphi(ndim,ntab,nbin,el,mass,hbin)
   integer(kind=i4) :: ndim,ntab,nbin(:)
   real(kind=r8) :: el,mass(:),hbin
   read (5,*) prm%v0
   read (5,*) prm%amu
   read (5,*) prm%cjas1
   read (5,*) prm%cjas2
   read (5,*) prm%djas
   read (5,*) prm%v0op
   read (5,*) prm%v0eq
   if (prm%cjas2.ne.0.0_r8) error stop 'setphi: equal-spin Jastrow (cjas(2) /= 0) is not supported'
   end subroutine setphi
end module potential

! The reference's random-number modules for host code (walker generation, dmcbcs.f90:237-301):
! the 48-bit linear congruential generators ran1 and ran2 (ran.f90:8-50) and the module
! stream with randn, gaussian (Box-Muller), setrn and savern (ran.f90:52-124).  The device
! kernels draw from the same generators with the same constants.
module random2
   implicit none
   integer, private, parameter :: i8=selected_int_kind(15)
   integer, private, parameter :: r8=selected_real_kind(15,9)
   integer(kind=i8), private, parameter :: m24=2_i8**24-1_i8, m48=2_i8**48-1_i8
contains
! s <- (a s + c) mod 2^48 with the product split in 24-bit halves so that no intermediate
! exceeds 63 bits: a s = a_lo s_lo + 2^24 (a_lo s_hi + a_hi s_lo) mod 2^48
   subroutine lcg48(s,a,c)
   integer(kind=i8) :: s,a,c
   integer(kind=i8) :: slo,shi,alo,ahi,mid
   slo=iand(s,m24)
   shi=iand(ishft(s,-24),m24)
   alo=iand(a,m24)
   ahi=iand(ishft(a,-24),m24)
   mid=iand(alo*shi+ahi*slo,m24)
   s=iand(alo*slo+ishft(mid,24)+c,m48)
   end subroutine lcg48

   subroutine ran1(rn,irn)
   real(kind=r8) :: rn
   integer(kind=i8) :: irn
   call lcg48(irn,44485709377909_i8,96309754297_i8)
   rn=real(ior(irn,1_i8),r8)*2.0_r8**(-48)
   end subroutine ran1

   subroutine ran2(rn,irn)
   real(kind=r8) :: rn
   integer(kind=i8) :: irn
   call lcg48(irn,34522712143931_i8,55789347517_i8)
   rn=real(ior(irn,1_i8),r8)*2.0_r8**(-48)
   end subroutine ran2
end module random2

module random
   use random2, only : ran1
   implicit none
   integer, private, parameter :: i4=selected_int_kind(9)
   integer, private, parameter :: i8=selected_int_kind(15)
   integer, private, parameter :: r8=selected_real_kind(15,9)
   integer(kind=i8), private, save :: stream=1_i8
contains
! n uniform variates from the module stream
   function randn(n)
   integer(kind=i4) :: n
   real(kind=r8) :: randn(n)
   integer(kind=i4) :: i
   do i=1,n
      call ran1(randn(i),stream)
   enddo
   end function randn

! n gaussian variates: 2 ceil(n/2) uniforms drawn first, then per pair (u1, u2)
! sin(2 pi u1) r and cos(2 pi u1) r with r = sqrt(-2 ln u2); an odd last variate takes the
! cosine of its pair
   function gaussian(n)
   integer(kind=i4) :: n
   real(kind=r8) :: gaussian(n)
   real(kind=r8) :: u(2*((n+1)/2)),arg,r
   integer(kind=i4) :: i
   u=randn(size(u))
   do i=1,(n+1)/2
      arg=2.0_r8*(4.0_r8*atan(1.0_r8))*u(2*i-1)
      r=sqrt(-2.0_r8*log(u(2*i)))
      if (2*i.le.n) then
         gaussian(2*i-1)=sin(arg)*r
         gaussian(2*i)=cos(arg)*r
      else
         gaussian(2*i-1)=cos(arg)*r
      endif
   enddo
   end function gaussian

   subroutine setrn(irnin)
   integer(kind=i8) :: irnin
   stream=iand(irnin,2_i8**48-1_i8)
   end subroutine setrn

   subroutine savern(irnout)
   integer(kind=i8) :: irnout
   irnout=stream
   end subroutine savern
end module random

module stack
   use engine
   implicit none
! the walker record of stack.f90:6-13, field for field (psigl, d2psig, dpsig and isg describe
! the guiding function, equal to psi here since a2 = 0; vext is 0 without an external
! potential; psil0 and psigl0 are the values read from a walker file; wt0 is declared
! real(kind=i4), i.e. default real, as the reference declares it)
   type walker
      real(kind=r8) :: psil=0.0_r8,d2psi=0.0_r8,v=0.0_r8,weight=1.0_r8,psigl=0.0_r8, &
         d2psig=0.0_r8,psil0=0.0_r8,psigl0=0.0_r8,vext=0.0_r8
      real(kind=r8), pointer :: x(:,:)=>null(),dpsi(:,:)=>null(),dpsig(:,:)=>null()
      integer(kind=i4) :: is=1,isg=1
      integer(kind=i8) :: irn=0
      real(kind=r8), pointer :: x0(:,:)=>null()
      real(kind=i4) :: wt0=1.0
   end type walker
   type(walker), save :: w1,w2
   integer(kind=i4), private, save :: npart,ndim,cap
! Host side of the reference's two stacks (stack.f90:15-17): walkers pushed onto stack i
! wait in host arrays until the next device call loads them (flush_stack); popping stack i
! when it is empty on the host downloads the device population into it first.  So a loop
! of dmcbcs.f90's form (pop stack istin, work on the walker, push stack istout) runs on the
! host, and the device holds the population between such loops.
   real(kind=r8), allocatable, target, private, save :: hx(:,:,:,:),hd(:,:,:,:),hsc(:,:,:)
   integer(kind=i8), allocatable, target, private, save :: hirn(:,:)
   integer(kind=i4), private, save :: nhost(2)=0
contains
   subroutine create(nst,nw,np,nd,euclid)
   integer(kind=i4) :: nst,nw,np,nd
   logical :: euclid
   npart=np
   ndim=nd
   cap=nw
   allocate(hx(nd,np,nw,2),hd(nd,np,nw,2),hsc(6,nw,2),hirn(nw,2))
   call walker_arrays(w1)
   call walker_arrays(w2)
   nhost=0
   end subroutine create

   subroutine walker_arrays(w)
   type(walker) :: w
   if (associated(w%x)) return
   allocate(w%x(ndim,npart),w%dpsi(ndim,npart),w%dpsig(ndim,npart),w%x0(ndim,npart))
   w%x=0.0_r8
   w%dpsi=0.0_r8
   w%dpsig=0.0_r8
   w%x0=0.0_r8
   end subroutine walker_arrays

! push (stack.f90:83-95): onto host stack i; it joins the device population at the next
! device call
   subroutine push(i,w)
   integer(kind=i4) :: i
   type(walker) :: w
   integer(kind=i4) :: k
   if (i.lt.1.or.i.gt.2) error stop 'push: stack number'
   if (nhost(i).ge.cap) error stop 'stack overflow'
   k=nhost(i)+1
   nhost(i)=k
   hx(:,:,k,i)=w%x
   hd(:,:,k,i)=w%dpsi
   hirn(k,i)=w%irn
   hsc(:,k,i)=(/w%weight,w%psil,w%d2psi,w%v,real(w%is,r8),real(w%wt0,r8)/)
   end subroutine push

! pop (stack.f90:97-112): from host stack i, which takes the device population when it is
! empty on the host
   subroutine pop(i,w,empty)
   integer(kind=i4) :: i
   type(walker) :: w
   logical :: empty
   integer(c_int32_t) :: nw
   integer(kind=i4) :: k
   if (i.lt.1.or.i.gt.2) error stop 'pop: stack number'
   if (nhost(i).eq.0.and.started) then
      call dmc_check(dmc_read_walkers(ctx,cap,nw,c_loc(hx(1,1,1,i)),c_loc(hd(1,1,1,i)), &
         c_loc(hsc(1,1,i)),c_loc(hirn(1,i))),'pop')
      nhost(i)=nw
      if (nw.gt.0) call dmc_check(dmc_load_walkers(ctx,0_c_int32_t,hx(:,:,:,i),hirn(:,i)),'pop')
   endif
   empty=nhost(i).eq.0
   if (empty) return
   k=nhost(i)
   call walker_arrays(w)
   w%x=hx(:,:,k,i)
   w%irn=hirn(k,i)
   w%dpsi=hd(:,:,k,i)
   w%weight=hsc(1,k,i)
   w%psil=hsc(2,k,i)
   w%d2psi=hsc(3,k,i)
   w%v=hsc(4,k,i)
   w%is=nint(hsc(5,k,i))
   w%wt0=real(hsc(6,k,i))
! guiding function = psi (a2 = 0); the device keeps no x0 (no euclidean estimators)
   w%dpsig=w%dpsi
   w%psigl=w%psil
   w%d2psig=w%d2psi
   w%isg=w%is
   w%psil0=w%psil
   w%psigl0=w%psil
   w%vext=0.0_r8
   w%x0=0.0_r8
   nhost(i)=k-1
   end subroutine pop

! walkers pushed on the host are loaded onto the device, stack 1 then stack 2 (positions and
! random states; the device evaluates hpsi and starts them with weight 1, wt0 = sign)
   subroutine flush_stack
   type(c_ptr) :: c
   integer(kind=i4) :: i,n
   integer(c_int32_t) :: nd
   real(kind=r8), allocatable, target :: xa(:,:,:)
   integer(kind=i8), allocatable, target :: ra(:)
   c=engine_ctx()
   n=sum(nhost)
   if (n.eq.0) return
   nd=0
   call dmc_check(dmc_population(c,nd),'load walkers')
   allocate(xa(ndim,npart,n+nd),ra(n+nd))
! walkers already on the device stay (downloaded and loaded again with the new ones)
   if (nd.gt.0) call dmc_check(dmc_read_walkers(c,int(n+nd,c_int32_t),nd,c_loc(xa), &
      c_null_ptr,c_null_ptr,c_loc(ra)),'load walkers')
   do i=1,2
      if (nhost(i).gt.0) then
         xa(:,:,nd+1:nd+nhost(i))=hx(:,:,1:nhost(i),i)
         ra(nd+1:nd+nhost(i))=hirn(1:nhost(i),i)
         nd=nd+nhost(i)
         nhost(i)=0
      endif
   enddo
   call dmc_check(dmc_load_walkers(c,nd,xa,ra),'load walkers')
   end subroutine flush_stack

! the resident walkers evaluated again with the current parameters (positions and random
! states kept; weight 1, wt0 = sign), as the set-up pass of dmcbcs.f90:314-353
   subroutine reload_stack
   integer(c_int32_t) :: nw
   call flush_stack
   call dmc_check(dmc_read_walkers(engine_ctx(),cap,nw,c_loc(hx),c_null_ptr,c_null_ptr, &
      c_loc(hirn)),'reload')
   call dmc_check(dmc_load_walkers(ctx,nw,hx(:,:,:,1),hirn(:,1)),'reload')
   end subroutine reload_stack

! walkers on stack i: those waiting on the host plus, for a stack empty on the host, the
! device population (which the next pop of that stack takes)
   function numstack(i)
   integer(kind=i4) :: numstack,i
   integer(c_int32_t) :: nw
   nw=0
   if (started) call dmc_check(dmc_population(ctx,nw),'numstack')
   numstack=nhost(i)
   if (nhost(i).eq.0) numstack=nw
   end function numstack
end module stack

module wavefunction
   use engine
   implicit none
   integer(kind=i4), private, save :: npart,norb
contains
! setpsi: nodes, a2 (wavefunction.f90:36-37), then the pairing deck (psislaterbcs.f90:47-68)
! and the beta deck (betafun.f90:19-21); returns the box side el
   subroutine setpsi(ndim,nbin,el,mass,ntab)
   integer(kind=i4) :: ndim,nbin(:),ntab
   real(kind=r8) :: el,mass(:)
   character(len=8) :: nodes
   real(kind=r8) :: a2,v0ex,v0exjas,jbeta,jgamma
   integer(kind=i4) :: nexup,nexdn,numq,i,nqex(3)
   read (5,'(a8)') nodes
   read (5,*) a2
   if (trim(adjustl(nodes)).ne.'bcs') error stop 'setpsi: only bcs nodes are supported'
   if (a2.ne.0.0_r8) error stop 'setpsi: a guiding function (a2 /= 0) is not supported'
   read (5,*) prm%rho
   read (5,*) nexup
   if (nexup.ne.0) error stop 'setpsi: unpaired up orbitals are not supported'
   read (5,*) nexdn
   if (nexdn.ne.0) error stop 'setpsi: unpaired down orbitals are not supported'
   read (5,*) norb
   if (norb.gt.16) error stop 'setpsi: more than 16 shells'
   prm%norb=norb
   prm%vorb=0.0_r8
   do i=1,norb
      read (5,*) prm%vorb(i)
   enddo
   read (5,*) nqex
   read (5,*) v0ex
   read (5,*) v0exjas
   read (5,*) numq
   read (5,*) jbeta
   read (5,*) jgamma
   if (any(nqex.ne.0).or.v0ex.ne.0.0_r8.or.v0exjas.ne.0.0_r8.or.numq.ne.0) &
      error stop 'setpsi: external potential and long-range Jastrow are not supported'
   read (5,*) prm%bbeta
   read (5,*) prm%dbeta
   read (5,*) prm%bfact
   if (prm%bfact.lt.0.0_r8) error stop 'setpsi: beta tables from file (bfact < 0) are not supported'
   npart=sum(nbin)
   prm%npart=npart
   prm%ntab=ntab
   prm%mass=mass(1)
   if (nbin(1).ne.nbin(2).or.mass(1).ne.mass(2)) error stop 'setpsi: unpolarised equal masses only'
   el=(npart/prm%rho)**(1.0_r8/ndim)
   end subroutine setpsi

! hpsi (wavefunction.f90:161-201) of one walker, on the device
   subroutine hpsi(w)
   use stack, only : walker
   type(walker) :: w
   real(kind=r8) :: psil(1),d2psi(1),v(1)
   integer(kind=i4) :: is(1)
   call dmc_check(dmc_eval(engine_ctx(),1_c_int32_t,w%x,psil,is,w%dpsi,d2psi,v),'hpsi')
   w%psil=psil(1)
   w%d2psi=d2psi(1)
   w%v=v(1)
   w%is=is(1)
! the guiding-function fields of wavefunction.f90:190-199 with a2 = 0, and no external
! potential
   w%psigl=w%psil
   w%d2psig=w%d2psi
   if (associated(w%dpsig)) w%dpsig=w%dpsi
   w%isg=w%is
   w%vext=0.0_r8
   end subroutine hpsi

! getderpsi (wavefunction.f90:203-229): d log psi / d(alpha_1..alpha_norb[, b, d]) at x
   subroutine getderpsi(x,dpsi)
   real(kind=r8) :: x(:,:),dpsi(:)
   real(kind=r8) :: d(norb+2)
   call dmc_check(dmc_adjoint(engine_ctx(),1_c_int32_t,x,d,c_null_ptr),'getderpsi')
   dpsi=d(1:size(dpsi))
   end subroutine getderpsi

! d log psi and d E_L with respect to the parameters for nw configurations x(3,npart,nw)
   subroutine getderelocal(nw,x,dlogpsi,delocal)
   integer(kind=i4) :: nw
   real(kind=r8), target :: x(:,:,:),dlogpsi(:,:),delocal(:,:)
   call dmc_check(dmc_adjoint(engine_ctx(),nw,x,dlogpsi,c_loc(delocal)),'getderelocal')
   end subroutine getderelocal

! setdetparam (wavefunction.f90:231-244): alpha_1..alpha_norb[, b, d]
   subroutine setdetparam(par)
   real(kind=r8) :: par(:)
   prm%vorb(1:norb)=par(1:norb)
   if (size(par).ge.norb+2) then
      prm%bbeta=par(norb+1)
      prm%dbeta=par(norb+2)
   endif
   if (started) call dmc_check(dmc_set_params(ctx,prm%vorb,prm%bbeta,prm%dbeta),'setdetparam')
   end subroutine setdetparam

   subroutine getdetparam(npar,par)
   integer(kind=i4) :: npar
   real(kind=r8) :: par(:)
   npar=norb
   par(1:norb)=prm%vorb(1:norb)
   if (size(par).ge.norb+2) then
      npar=norb+2
      par(norb+1)=prm%bbeta
      par(norb+2)=prm%dbeta
   endif
   end subroutine getdetparam

! move1 (wavefunction.f90:393-404): psi(x with particle i at xnew)/psi(x)
   subroutine move1(i,xnew,w,psirat)
   use stack, only : walker
   integer(kind=i4) :: i
   real(kind=r8) :: xnew(:),psirat
   type(walker) :: w
   integer(kind=i4) :: ip(1)
   real(kind=r8) :: r(1)
   ip(1)=i-1
   call dmc_check(dmc_ratio1(engine_ctx(),1_c_int32_t,w%x,ip,xnew,r),'move1')
   psirat=r(1)
   end subroutine move1
end module wavefunction

module step
   use engine
   use stack, only : flush_stack
   implicit none
   integer(kind=i4), private, save :: idmc
contains
! setstep (step.f90:11-41).  The engine derives sigma, the drift cut and the internal trial
! energy from the deck values itself; the scaled etrial and eunit passed here are not used.
   subroutine setstep(idmcin,ndim,npart,nbin,hbar,dt,etrial,el,mass,wtcut,dcut,eunit,fn)
   integer(kind=i4) :: idmcin,ndim,npart,nbin(:)
   real(kind=r8) :: hbar,dt,etrial,el,mass(:),wtcut,dcut,eunit
   logical :: fn
   idmc=idmcin
   if (idmc.ne.1.and.idmc.ne.2.and.idmc.ne.0.and.idmc.ne.-1) &
      error stop 'setstep: the engine runs DMC (idmc = 1, 2) and VMC (idmc = 0, -1)'
   prm%hbar=hbar
   prm%dt=dt
   prm%wtcut=wtcut
   prm%dcut=dcut
   prm%fn=merge(1,0,fn)
   end subroutine setstep

! step1 (step.f90:43-285): the whole population is propagated on the device
   subroutine step1(istin,istout,opcomp)
   integer(kind=i4) :: istin,istout
   logical :: opcomp
   call flush_stack
   call dmc_check(dmc_propagate(engine_ctx(),merge(1,0,opcomp)),'step1')
   end subroutine step1

! branch (step.f90:287-315)
   subroutine branch(istin,istout,factor)
   integer(kind=i4) :: istin,istout
   real(kind=r8) :: factor
   integer(c_int32_t) :: nt
   call flush_stack
   call dmc_check(dmc_branch(engine_ctx(),factor,nt),'branch')
   end subroutine branch

! writewtisto (step.f90:317-334): histogram of branch multiplicities
   subroutine writewtisto
   integer(c_int64_t) :: hist(500)
   integer(kind=i4) :: i
   call dmc_check(dmc_wtisto(engine_ctx(),hist),'writewtisto')
   open (unit=33,file='wtisto',status='replace')
   do i=1,500
      if (hist(i).gt.0) write (33,'(i5,i14)') i-1,hist(i)
   enddo
   close(33)
   end subroutine writewtisto
end module step

module estimator
   use engine
   implicit none
   integer(kind=i4), private, save :: nest=0
   character(len=30), allocatable, private, save :: label(:)
   real(kind=r8), allocatable, private, save :: valnow(:),errnow(:)
contains
   subroutine setestnum(n)
   integer(kind=i4) :: n
   nest=n
   allocate(label(n),valnow(n),errnow(n))
   label=' '
   valnow=0.0_r8
   errnow=0.0_r8
   end subroutine setestnum

   subroutine addest(i,l)
   integer(kind=i4) :: i
   character(len=*) :: l
   label(i)=l
   end subroutine addest

! zerest: discard the sums accumulated so far (end of equilibration)
   subroutine zerest
   real(c_double) :: est(3,14)
   call dmc_check(dmc_estimators(engine_ctx(),est,1_c_int32_t),'zerest')
   end subroutine zerest

! update (estimator.f90:84-112): block value and its error from the device block sums
! (sum val*wt, sum val^2*wt, sum wt), then the sums are cleared
   subroutine update
   real(c_double) :: est(3,14)
   integer(kind=i4) :: i
   real(kind=r8) :: v2
   call dmc_check(dmc_estimators(engine_ctx(),est,1_c_int32_t),'update')
   do i=1,nest
      valnow(i)=est(1,i)/est(3,i)
      v2=est(2,i)/est(3,i)
      errnow(i)=sqrt(abs(v2-valnow(i)**2)/est(3,i))
   enddo
   end subroutine update

! resstring (estimator.f90:135-149): one line per estimator, block value +- block error
   function resstring(tau)
   real(kind=r8) :: tau
   character(len=90), dimension(nest) :: resstring
   integer(kind=i4) :: i
   do i=1,nest
      write (resstring(i),'(a30,t30,1x,1p,e15.8,e19.10,'' +- '',e19.10)') &
         label(i),tau,valnow(i),errnow(i)
   enddo
   end function resstring
end module estimator

module optimizer
   use engine
   implicit none
   real(kind=r8), private, save :: delta0,eps
   integer(kind=i4), private, save :: nparm
   logical, private, save :: iopt(2)
! sums of the per-walker calder(w, ene0), in the layout of dmc_opt_sums: n, sum E, sum E^2,
! sum O_p, sum O_p E, sum O_p O_q (p, q over alpha_1..alpha_norb, b, d)
   real(kind=r8), allocatable, private, save :: hsums(:)
! calder(w, ene0) for one walker (optimizer.f90:90-112) and calder with no arguments for
! every resident walker at once on the device
   interface calder
      module procedure calder_walker, calder_population
   end interface calder
contains
! setupopt (optimizer.f90:14-69): optiter, optimise Jastrow, optimise nodes, delta, eps.
! The parameters are the determinant's alpha_1..alpha_norb (getdetparam); the cosh Jastrow
! has none.
   subroutine setupopt(npart,optiter)
   integer(kind=i4) :: npart,optiter
   read (5,*) optiter
   read (5,*) iopt(1)
   read (5,*) iopt(2)
   read (5,*) delta0
   read (5,*) eps
   nparm=0
   if (iopt(2)) nparm=prm%norb
   write (6,'(''iterations for the optimization'',t40,i10)') optiter
   write (6,'(''optimize Jastrow'',t40,l10)') iopt(1)
   write (6,'(''optimize nodes'',t40,l10)') iopt(2)
   write (6,'(''delta ='',t40,f10.5)') delta0
   write (6,'(''epsilon ='',t40,f10.5)') eps
   write (6,'(''number of parameters ='',t40,i10)') nparm
   end subroutine setupopt

! calder (optimizer.f90:90-112) for every resident walker at once: the device computes
! d log psi with the adjoint kernel and accumulates the sums (E = each walker's stored
! E_L / (npart eunit))
   subroutine calder_population
   call dmc_check(dmc_opt_accumulate(engine_ctx()),'calder')
   end subroutine calder_population

! calder(w, ene0) (optimizer.f90:90-112) for one walker: d log psi at w%x from the adjoint
! kernel, added with the caller's energy ene0 to the host sums
   subroutine calder_walker(w,ene0)
   use stack, only : walker
   type(walker) :: w
   real(kind=r8) :: ene0
   real(kind=r8) :: d(prm%norb+2)
   integer(kind=i4) :: np,i,j
   np=prm%norb+2
   if (.not.allocated(hsums)) then
      allocate(hsums(3+2*np+np*np))
      hsums=0.0_r8
   endif
   call dmc_check(dmc_adjoint(engine_ctx(),1_c_int32_t,w%x,d,c_null_ptr),'calder')
   hsums(1)=hsums(1)+1.0_r8
   hsums(2)=hsums(2)+ene0
   hsums(3)=hsums(3)+ene0**2
   do i=1,np
      hsums(3+i)=hsums(3+i)+d(i)
      hsums(3+np+i)=hsums(3+np+i)+d(i)*ene0
      do j=1,np
         hsums(3+2*np+(i-1)*np+j)=hsums(3+2*np+(i-1)*np+j)+d(i)*d(j)
      enddo
   enddo
   end subroutine calder_walker

! updatepar (optimizer.f90:114-169): f = -2<O E> + 2<O><E>, S = <O O> - <O><O> + eps,
! parameters += delta S^-1 f; the new parameters go to the wavefunction (setdetparam)
   subroutine updatepar(ene)
   real(kind=r8) :: ene
   real(c_double) :: sums(3+2*(prm%norb+2)+(prm%norb+2)**2)
   real(kind=r8) :: ovec(nparm),ovecw(nparm),pmat(nparm,nparm),fvec(nparm),dpar(nparm)
   real(kind=r8) :: n,esum,esum2,eerr
   integer(kind=i4) :: i,j,np
   np=prm%norb+2
   call dmc_check(dmc_opt_sums(engine_ctx(),sums,1_c_int32_t),'updatepar')
! the sums of both forms of calder
   if (allocated(hsums)) then
      sums=sums+hsums
      hsums=0.0_r8
   endif
   n=sums(1)
   esum=sums(2)/n
   esum2=sums(3)/n
   eerr=sqrt(abs(esum**2-esum2)/n)
   ene=esum
   if (nparm.eq.0) return
   do i=1,nparm
      ovec(i)=sums(3+i)/n
      ovecw(i)=sums(3+np+i)/n
   enddo
   do i=1,nparm
      do j=1,nparm
         pmat(i,j)=sums(3+2*np+(i-1)*np+j)/n-ovec(i)*ovec(j)
      enddo
      pmat(i,i)=pmat(i,i)+eps
   enddo
   fvec=-2.0_r8*ovecw+2.0_r8*ovec*esum
   call solve(nparm,pmat,fvec,dpar)
   prm%vorb(1:nparm)=prm%vorb(1:nparm)+delta0*dpar
   call dmc_check(dmc_set_params(engine_ctx(),prm%vorb,prm%bbeta,prm%dbeta),'updatepar')
   write (6,*) 'eave = ',ene,eerr
   write (6,'(''new parameters: '')')
   do i=1,nparm
      write (6,'(f30.20)') prm%vorb(i)
   enddo
   end subroutine updatepar

! x = a^-1 b by Gauss-Jordan elimination with partial pivoting (a is overwritten)
   subroutine solve(n,a,b,x)
   integer(kind=i4) :: n
   real(kind=r8) :: a(n,n),b(n),x(n)
   integer(kind=i4) :: k,p,i
   real(kind=r8) :: f,t(n),tb
   x=b
   do k=1,n
      p=k-1+maxloc(abs(a(k:n,k)),1)
      if (a(p,k).eq.0.0_r8) error stop 'updatepar: singular parameter matrix'
      if (p.ne.k) then
         t=a(k,:)
         a(k,:)=a(p,:)
         a(p,:)=t
         tb=x(k)
         x(k)=x(p)
         x(p)=tb
      endif
      f=1.0_r8/a(k,k)
      a(k,:)=a(k,:)*f
      x(k)=x(k)*f
      do i=1,n
         if (i.ne.k) then
            f=a(i,k)
            a(i,:)=a(i,:)-f*a(k,:)
            x(i)=x(i)-f*x(k)
         endif
      enddo
   enddo
   end subroutine solve
end module optimizer

! g(r) and the density profile (gofr.f90:14-177).  The reference adds one walker per call
! (addgofr); here addgofr_population adds every walker of the device population with the
! weight of addoperators in one call of dmc_gofr.  Block normalisation and the tables are the
! reference's host arithmetic on one process (its addall is the identity there).
module gofr
   use engine
   implicit none
   private
   public :: setupgofr, zerogofr, addgofr_population, updategofr, writegofr
   integer(kind=i4), save :: npart, ntab, ndim
   real(kind=r8), save :: el, dr, wsblk, totalw, kf
   real(kind=r8), save, allocatable :: gnow(:,:), gerr(:,:), g(:,:), ge(:,:)
   real(kind=r8), save, allocatable :: dnow(:,:), derr(:,:), d(:,:), de(:,:)
contains
   subroutine setupgofr(ndin,ntabin,nbinin,elin,kfin)
   integer(kind=i4) :: ndin,ntabin,nbinin(:)
   real(kind=r8) :: elin,kfin
   if (ndin.ne.3) error stop 'setupgofr: three dimensions only'
   ndim=ndin
   ntab=ntabin
   npart=sum(nbinin)
   el=elin
   kf=kfin
   allocate(gnow(3,0:ntab),gerr(3,0:ntab),g(3,0:ntab),ge(3,0:ntab))
   allocate(dnow(6,-ntab:ntab),derr(6,-ntab:ntab),d(6,-ntab:ntab),de(6,-ntab:ntab))
   dr=0.5_r8*el/ntab
   call zerogofr
   end subroutine setupgofr

   subroutine zerogofr
   g=0.0_r8
   gnow=0.0_r8
   gerr=0.0_r8
   ge=0.0_r8
   d=0.0_r8
   dnow=0.0_r8
   derr=0.0_r8
   de=0.0_r8
   wsblk=0.0_r8
   end subroutine zerogofr

   subroutine addgofr_population
   call dmc_check(dmc_gofr(engine_ctx(),ntab,gnow,gerr,dnow,derr,wsblk),'dmc_gofr')
   end subroutine addgofr_population

   subroutine updategofr
   real(kind=r8) :: norm
   totalw=wsblk
   norm=1.0_r8/wsblk
   g=gnow*norm
   ge=sqrt(abs(gerr*norm-g**2)/wsblk)
   d=dnow*norm
   de=sqrt(abs(derr*norm-d**2)/wsblk)
   wsblk=0.0_r8
   gnow=0.0_r8
   gerr=0.0_r8
   dnow=0.0_r8
   derr=0.0_r8
   end subroutine updategofr

   subroutine writegofr(filename,tau)
   character(len=*) :: filename
   real(kind=r8) :: tau,pi,r0,r1,vol,rho,facl
   integer(kind=i4) :: i
   pi=4.0_r8*atan(1.0_r8)
   rho=npart/el**ndim
   open (unit=77,file=filename,position='append')
   write (77,'(''# r*kF, g(r), err'')')
   write (77,'(''# total weight = '',e15.7)') totalw
   write (77,'(''# tau = '',e15.7)') tau
   do i=0,ntab-1
      r0=i*dr
      r1=(i+1)*dr
      vol=4.0_r8*pi*(r1**3-r0**3)/3.0_r8
      facl=1.0_r8/(rho*npart*vol)
      write (77,'(7e15.7)') (i+0.5_r8)*dr,g(1,i)*facl,ge(1,i)*facl,g(2,i)*facl,ge(2,i)*facl, &
         g(3,i)*facl,ge(3,i)*facl
   enddo
   write (77,*) ''
   write (77,*) ''
   close (77)
   open (unit=77,file='rho.out',position='append')
   write (77,'(''# r, xrhoup, err, xrhodn, err, yrhoup, err, yrhodn, err, zrhoup, err, zrhodn, err'')')
   write (77,'(''# total weight = '',e15.7)') totalw
   write (77,'(''# tau = '',e15.7)') tau
   facl=1.0_r8/npart
   do i=-ntab,ntab
      write (77,'(13e15.7)') i*dr,d(1,i)*facl,de(1,i)*facl,d(4,i)*facl,de(4,i)*facl, &
         d(2,i)*facl,de(2,i)*facl,d(5,i)*facl,de(5,i)*facl,d(3,i)*facl,de(3,i)*facl, &
         d(6,i)*facl,de(6,i)*facl
   enddo
   write (77,*) ''
   write (77,*) ''
   close (77)
   end subroutine writegofr
end module gofr

! S(q) (sofq.f90:14-119): the q vectors of setupk come from the engine (dmc_sofq_kvec), the
! sums of every walker from one call of dmc_sofq.  The deck is read before the engine's
! context exists (setstep comes later), so the q vectors and the accumulators are set up at
! the first zero or add.
module sofq
   use engine
   implicit none
   private
   public :: setupsofq, zerosofq, addsofq_population, updatesofq, writesofq
   integer(kind=i4), save :: npart, ntabq, nqtot
   logical, save :: ready=.false.
   integer(c_int32_t), save, allocatable, target :: numsh(:)
   real(kind=r8), save :: wsblk, totalw, kf
   real(c_double), save, allocatable, target :: ak(:,:)
   real(kind=r8), save, allocatable :: snow(:,:), s(:,:), serr(:,:), se(:,:)
contains
   subroutine setupsofq(ndin,nbinin,el,ntabqin,kfin)
   integer(kind=i4) :: ndin,nbinin(:),ntabqin
   real(kind=r8) :: el,kfin
   if (ndin.ne.3) error stop 'setupsofq: three dimensions only'
   npart=sum(nbinin)
   ntabq=ntabqin
   kf=kfin
   end subroutine setupsofq

! setupk (kshell.f90:6-18) through the engine, and the accumulators
   subroutine kvectors
   integer(c_int32_t) :: nq
   if (ready) return
   allocate(numsh(ntabq))
   call dmc_check(dmc_sofq_kvec(engine_ctx(),ntabq,0,nq,c_null_ptr,c_null_ptr),'dmc_sofq_kvec')
   nqtot=nq
   allocate(ak(3,nqtot))
   call dmc_check(dmc_sofq_kvec(engine_ctx(),ntabq,nq,nq,c_loc(numsh),c_loc(ak)), &
      'dmc_sofq_kvec')
   allocate(snow(3,nqtot),s(3,nqtot),serr(3,nqtot),se(3,nqtot))
   s=0.0_r8
   snow=0.0_r8
   serr=0.0_r8
   se=0.0_r8
   wsblk=0.0_r8
   ready=.true.
   end subroutine kvectors

   subroutine zerosofq
   call kvectors
   s=0.0_r8
   snow=0.0_r8
   serr=0.0_r8
   se=0.0_r8
   wsblk=0.0_r8
   end subroutine zerosofq

   subroutine addsofq_population
   call kvectors
   call dmc_check(dmc_sofq(engine_ctx(),ntabq,snow,serr,wsblk),'dmc_sofq')
   end subroutine addsofq_population

   subroutine updatesofq
   real(kind=r8) :: norm
   totalw=wsblk
   norm=1.0_r8/wsblk
   s=snow*norm
   se=sqrt(abs(serr*norm-s**2)/wsblk)
   wsblk=0.0_r8
   snow=0.0_r8
   serr=0.0_r8
   end subroutine updatesofq

   subroutine writesofq(filename,tau)
   character(len=*) :: filename
   real(kind=r8) :: tau,ave(3),err(3),q
   integer(kind=i4) :: iq,j,k
   open (unit=79,file=filename,position='append')
   write (79,'(''# q/kF, s(q), err'')')
   write (79,'(''# total weight = '',e15.7)') totalw
   write (79,'(''# tau = '',e15.7)') tau
   k=0
   do iq=1,ntabq
      ave=0.0_r8
      err=0.0_r8
      do j=1,numsh(iq)
         k=k+1
         ave=ave+s(:,k)
         err=err+se(:,k)
      enddo
      ave=ave/numsh(iq)/npart
      err=err/numsh(iq)/npart
      q=sqrt(sum(ak(:,k)**2))
      write (79,'(7e15.7)') q/kf,ave(1),err(1),ave(2),err(2),ave(3),err(3)
   enddo
   write (79,*) ' '
   write (79,*) ' '
   close (79)
   end subroutine writesofq
end module sofq

! n(k) (momentum3d.f90:13-190): the shells of setupenk come from the engine
! (dmc_nofk_shells) at the first zero or add, the sums of every walker from one call of
! dmc_nofk.
module pdistribution
   use engine
   implicit none
   private
   public :: setupenk, zeroenk, addenk_population, updateenk, writeenk
   integer(kind=i4), save :: ntab, nsh
   logical, save :: ready=.false.
   real(kind=r8), save :: wsblk, totalw
   real(c_double), save, allocatable, target :: kmag(:)
   integer(c_int32_t), save, allocatable, target :: nk(:)
   real(kind=r8), save, allocatable :: enknow(:), enkerr(:), enk(:), enke(:)
contains
   subroutine setupenk(ndim,ntabin,elin,kfin)
   integer(kind=i4) :: ndim,ntabin
   real(kind=r8) :: elin,kfin
   if (ndim.ne.3) error stop 'setupenk: three dimensions only'
   ntab=ntabin
   end subroutine setupenk

   subroutine shells
   integer(c_int32_t) :: n
   if (ready) return
   call dmc_check(dmc_nofk_shells(engine_ctx(),ntab,0,n,c_null_ptr,c_null_ptr), &
      'dmc_nofk_shells')
   nsh=n
   allocate(kmag(nsh),nk(nsh),enknow(nsh),enkerr(nsh),enk(nsh),enke(nsh))
   call dmc_check(dmc_nofk_shells(engine_ctx(),ntab,n,n,c_loc(kmag),c_loc(nk)), &
      'dmc_nofk_shells')
   enknow=0.0_r8
   enkerr=0.0_r8
   enk=0.0_r8
   enke=0.0_r8
   wsblk=0.0_r8
   ready=.true.
   end subroutine shells

   subroutine zeroenk
   call shells
   enk=0.0_r8
   enknow=0.0_r8
   enkerr=0.0_r8
   enke=0.0_r8
   wsblk=0.0_r8
   end subroutine zeroenk

   subroutine addenk_population
   call shells
   call dmc_check(dmc_nofk(engine_ctx(),ntab,enknow,enkerr,wsblk),'dmc_nofk')
   end subroutine addenk_population

   subroutine updateenk
   real(kind=r8) :: norm
   totalw=wsblk
   norm=1.0_r8/wsblk
   enk=enknow*norm
   enke=sqrt(abs(enkerr*norm-enk**2)/wsblk)
   wsblk=0.0_r8
   enknow=0.0_r8
   enkerr=0.0_r8
   end subroutine updateenk

   subroutine writeenk(filenamek,tau)
   character(len=*) :: filenamek
   real(kind=r8) :: tau
   integer(kind=i4) :: ir
   open (unit=78,file=filenamek,position='append')
   write (78,'(''# k/kF, n(k), err'')')
   write (78,'(''# total weight = '',e15.7)') totalw
   write (78,'(''# tau = '',e15.7)') tau
   do ir=1,nsh
      write (78,'(6e15.7)') kmag(ir),enk(ir),enke(ir)
   enddo
   write (78,*) ''
   write (78,*) ''
   close (78)
   end subroutine writeenk
end module pdistribution

! operators.f90:7-140 for a periodic box: the four table sizes of the deck; g(r), S(q) and
! n(k) are computed, the euclidean density response stops the run.  The walker loops of
! dmcbcs.f90:385-390,465-471 (pop, addoperators, push) become one call per observable over
! the device population.
module operators
   use engine
   use gofr
   use sofq
   use pdistribution
   implicit none
   private
   public :: setoperators, zerooperators, addoperators, updateoperators, writeoperators
   logical, save :: igofr=.false., isofq=.false., ienk=.false.
contains
   subroutine setoperators(ndim,nbin,el,oper,kunit)
   integer(kind=i4) :: ndim,nbin(:)
   real(kind=r8) :: el,kunit
   logical :: oper
   integer(kind=i4) :: ntabg,ntabq,ntabenk,ntabqeuc
   if (el.le.0.0_r8) error stop 'setoperators: periodic boxes only'
   read (5,*) ntabg
   read (5,*) ntabq
   read (5,*) ntabenk
   read (5,*) ntabqeuc
   if (ntabqeuc.gt.0) &
      error stop 'setoperators: the euclidean density response is not supported'
   igofr=ntabg.gt.0
   isofq=ntabq.gt.0
   ienk=ntabenk.gt.0
   if (igofr) call setupgofr(ndim,ntabg,nbin,el,kunit)
   if (isofq) call setupsofq(ndim,nbin,el,ntabq,kunit)
   if (ienk) call setupenk(ndim,ntabenk,el,kunit)
   oper=igofr.or.isofq.or.ienk
   end subroutine setoperators

   subroutine zerooperators
   if (igofr) call zerogofr
   if (isofq) call zerosofq
   if (ienk) call zeroenk
   end subroutine zerooperators

   subroutine addoperators
   if (igofr) call addgofr_population
   if (isofq) call addsofq_population
   if (ienk) call addenk_population
   end subroutine addoperators

   subroutine updateoperators
   if (igofr) call updategofr
   if (isofq) call updatesofq
   if (ienk) call updateenk
   end subroutine updateoperators

   subroutine writeoperators(tau)
   real(kind=r8) :: tau
   if (igofr) call writegofr('gofr.out',tau)
   if (isofq) call writesofq('sofq.out',tau)
   if (ienk) call writeenk('nofk.out',tau)
   end subroutine writeoperators
end module operators
