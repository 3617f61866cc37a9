! iso_c_binding interfaces to the C-ABI of include/dmc_gpu.h (libdmc_gpu.so).
! One interface per entry point; argument meaning as documented in the header.
module dmc_gpu
   use iso_c_binding
   implicit none
   integer(c_int), parameter :: DMC_OK=0, DMC_ERR_ARG=1, DMC_ERR_HIP=2, DMC_ERR_OVERFLOW=3, &
      DMC_ERR_SETUP=4, DMC_ERR_NODEVICE=5
   integer(c_int32_t), parameter :: DMC_FP64=0, DMC_MIXED=1
   integer(c_int32_t), parameter :: DMC_ALLPARTICLE=0, DMC_SINGLEPARTICLE=1, DMC_VMC=2, &
      DMC_METROPOLIS=3

   type, bind(c) :: dmc_params
      integer(c_int32_t) :: npart, norb, ntab, fn, units, nwalk, capacity, precision, &
                            movetype, krefresh
      real(c_double) :: rho, hbar, mass, dt, etrial, dcut, wtcut, popup, popdn
      real(c_double) :: bbeta, dbeta, bfact, v0, amu, cjas1, cjas2, djas, v0op, v0eq
      real(c_double) :: vorb(16)
   end type dmc_params

   interface
      integer(c_int) function dmc_create(p, ctx) bind(c, name='dmc_create')
         import :: c_int, c_ptr, dmc_params
         type(dmc_params), intent(in) :: p
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function dmc_destroy(ctx) bind(c, name='dmc_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function dmc_info(ctx, out) bind(c, name='dmc_info')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: out(8)
      end function
      type(c_ptr) function dmc_last_error() bind(c, name='dmc_last_error')
         import :: c_ptr
      end function
      integer(c_int) function dmc_set_params(ctx, vorb, bbeta, dbeta) bind(c, name='dmc_set_params')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: vorb(*)
         real(c_double), value :: bbeta, dbeta
      end function
      integer(c_int) function dmc_load_walkers(ctx, nw, x, irn) bind(c, name='dmc_load_walkers')
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nw
         real(c_double), intent(in) :: x(*)
         integer(c_int64_t), intent(in) :: irn(*)
      end function
      integer(c_int) function dmc_init_walkers(ctx, seed, nwalk) bind(c, name='dmc_init_walkers')
         import :: c_int, c_ptr, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: seed
         integer(c_int32_t), value :: nwalk
      end function
      integer(c_int) function dmc_read_walkers(ctx, cap, nw, x, dpsi, sc, irn) &
            bind(c, name='dmc_read_walkers')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: cap
         integer(c_int32_t), intent(out) :: nw
         type(c_ptr), value :: x, dpsi, sc, irn
      end function
      integer(c_int) function dmc_population(ctx, nw) bind(c, name='dmc_population')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), intent(out) :: nw
      end function
      integer(c_int) function dmc_eval(ctx, nw, x, psil, is, dpsi, d2psi, v) bind(c, name='dmc_eval')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nw
         real(c_double), intent(in) :: x(*)
         real(c_double), intent(out) :: psil(*), dpsi(*), d2psi(*), v(*)
         integer(c_int32_t), intent(out) :: is(*)
      end function
      integer(c_int) function dmc_ratio1(ctx, nw, x, ip, xnew, ratio) bind(c, name='dmc_ratio1')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nw
         real(c_double), intent(in) :: x(*), xnew(*)
         integer(c_int32_t), intent(in) :: ip(*)
         real(c_double), intent(out) :: ratio(*)
      end function
      integer(c_int) function dmc_propagate(ctx, opcomp) bind(c, name='dmc_propagate')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: opcomp
      end function
      integer(c_int) function dmc_branch(ctx, factor, nt) bind(c, name='dmc_branch')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: factor
         integer(c_int32_t), intent(out) :: nt
      end function
      integer(c_int) function dmc_step(ctx, opcomp, nt) bind(c, name='dmc_step')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: opcomp
         integer(c_int32_t), intent(out) :: nt(2)
      end function
      integer(c_int) function dmc_estimators(ctx, est, reset) bind(c, name='dmc_estimators')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: est(3,14)
         integer(c_int32_t), value :: reset
      end function
      integer(c_int) function dmc_wtisto(ctx, hist) bind(c, name='dmc_wtisto')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: hist(500)
      end function
      integer(c_int) function dmc_adjoint(ctx, nw, x, dlogpsi, delocal) bind(c, name='dmc_adjoint')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nw
         real(c_double), intent(in) :: x(*)
         real(c_double), intent(out) :: dlogpsi(*)
         type(c_ptr), value :: delocal
      end function
      integer(c_int) function dmc_opt_accumulate(ctx) bind(c, name='dmc_opt_accumulate')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function dmc_move_trace(ctx, enable) bind(c, name='dmc_move_trace')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: enable
      end function
      integer(c_int) function dmc_read_move_trace(ctx, cap, nw, rec) &
            bind(c, name='dmc_read_move_trace')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: cap
         integer(c_int32_t), intent(out) :: nw
         real(c_double), intent(out) :: rec(*)
      end function
      integer(c_int) function dmc_opt_accumulate_noalloc(ctx) &
            bind(c, name='dmc_opt_accumulate_noalloc')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function dmc_opt_sums(ctx, sums, reset) bind(c, name='dmc_opt_sums')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: sums(*)
         integer(c_int32_t), value :: reset
      end function
      integer(c_int) function dmc_gofr(ctx, ntab, gnow, gerr, dnow, derr, wsblk) &
            bind(c, name='dmc_gofr')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ntab
         real(c_double), intent(inout) :: gnow(*), gerr(*), dnow(*), derr(*), wsblk
      end function
      integer(c_int) function dmc_sofq_kvec(ctx, ntabq, cap, nqtot, numsh, ak) &
            bind(c, name='dmc_sofq_kvec')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ntabq, cap
         integer(c_int32_t), intent(out) :: nqtot
         type(c_ptr), value :: numsh, ak
      end function
      integer(c_int) function dmc_sofq(ctx, ntabq, snow, serr, wsblk) bind(c, name='dmc_sofq')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ntabq
         real(c_double), intent(inout) :: snow(*), serr(*), wsblk
      end function
      integer(c_int) function dmc_nofk_shells(ctx, ntab, cap, nsh, kmag, nk) &
            bind(c, name='dmc_nofk_shells')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ntab, cap
         integer(c_int32_t), intent(out) :: nsh
         type(c_ptr), value :: kmag, nk
      end function
      integer(c_int) function dmc_nofk(ctx, ntab, enknow, enkerr, wsblk) bind(c, name='dmc_nofk')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: ntab
         real(c_double), intent(inout) :: enknow(*), enkerr(*), wsblk
      end function
      integer(c_int) function dmc_sync(ctx) bind(c, name='dmc_sync')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
   end interface

contains

! the message of the last failed call, as a Fortran string
   function dmc_error_message() result(msg)
   character(len=256) :: msg
   character(kind=c_char), pointer :: cs(:)
   integer :: i
   msg=' '
   call c_f_pointer(dmc_last_error(), cs, [256])
   do i=1,256
      if (cs(i).eq.c_null_char) exit
      msg(i:i)=cs(i)
   enddo
   end function dmc_error_message

! stop with the library's message on a failed call (the reference stops on errors)
   subroutine dmc_check(ierr, what)
   integer(c_int), intent(in) :: ierr
   character(len=*), intent(in) :: what
   if (ierr.ne.DMC_OK) then
      write (6,'(a,'' failed ('',i0,''): '',a)') what, ierr, trim(dmc_error_message())
      stop 2
   endif
   end subroutine dmc_check
end module dmc_gpu
