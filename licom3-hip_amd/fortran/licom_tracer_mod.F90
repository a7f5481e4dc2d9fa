!-----------------------------------------------------------------------
! licom_tracer_mod -- ISO_C_BINDING interface to liblicom_tracer.so, the
! gfx950 tracer step.  Declares the C structs of include/licom_tracer.h as
! bind(C) derived types and the C functions as bind(C) interfaces, so a
! Fortran caller (licom_drv.F90 / stepon) can call the library directly,
! passing its own arrays by reference.
!
! Layout is LICOM's: (imt,jmt,km[,ntra]) with i fastest, two ghost cells,
! atb(imt,jmt,0:km,ntra), ws(imt,jmt,km+1) -- no transposition anywhere.
! Grid pointers (licom_grid) are host arrays; fields and diagnostics are
! device arrays (hipMalloc'd by the caller, e.g. the d_* globals that
! cuda_init.cpp:372-707 allocates).
!
! Not compiled in this repository (no Fortran compiler in the build image);
! it follows the C header field for field.
!-----------------------------------------------------------------------
module licom_tracer_mod
   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: LICOM_OK = 0
   integer(c_int), parameter, public :: LICOM_STEP_CONVADJ = 1
   integer(c_int), parameter, public :: LICOM_STEP_SIDE_EFFECTS = 2
   integer(c_int), parameter, public :: LICOM_STEP_VTL_SCRATCH = 4
   integer(c_int), parameter, public :: LICOM_STEP_STAGED = 8
   integer(c_int), parameter, public :: LICOM_STEP_ICESNOW = 16
   integer(c_int), parameter, public :: LICOM_HMIX_DEL2 = 2
   integer(c_int), parameter, public :: LICOM_HMIX_DEL4 = 4
   integer(c_int), parameter, public :: LICOM_ADV_CENTERED = 0
   integer(c_int), parameter, public :: LICOM_ADV_FLUX = 1
   integer(c_int), parameter, public :: LICOM_ADV_TSPAS = 2

   type, bind(C), public :: licom_grid
      integer(c_int) :: nx, ny, km, imt_global
      type(c_ptr) :: kmt = c_null_ptr          ! int32 (imt,jmt)
      type(c_ptr) :: zkp = c_null_ptr          ! (km+1)
      type(c_ptr) :: pen = c_null_ptr          ! (km-1) or null
      type(c_ptr) :: htw = c_null_ptr, hts = c_null_ptr
      type(c_ptr) :: hun = c_null_ptr, hue = c_null_ptr
      type(c_ptr) :: tarea = c_null_ptr
      type(c_ptr) :: au0 = c_null_ptr, aus = c_null_ptr, auw = c_null_ptr, ausw = c_null_ptr
      type(c_ptr) :: to = c_null_ptr, so = c_null_ptr
      type(c_ptr) :: eosc = c_null_ptr         ! C(km,9)
      type(c_ptr) :: dxu = c_null_ptr, dyu = c_null_ptr   ! (imt,jmt), flux scheme only
   end type licom_grid

   type, bind(C), public :: licom_params
      real(c_double) :: dts, onbc, oncc, dwndmix, gamma, od0, od0cp, ah
      integer(c_int) :: hmix, adv, boundary_restore, sssnorm
      integer(c_int) :: nss                    ! pconst_mod nss (centered diagnostics)
   end type licom_params

   type, bind(C), public :: licom_fields
      type(c_ptr) :: at = c_null_ptr, atb = c_null_ptr, akt = c_null_ptr
      type(c_ptr) :: utf = c_null_ptr, vtf = c_null_ptr, utl = c_null_ptr, vtl = c_null_ptr
      type(c_ptr) :: ws = c_null_ptr, h0f = c_null_ptr, h0l = c_null_ptr
      type(c_ptr) :: swv = c_null_ptr, nswv = c_null_ptr, sst = c_null_ptr
      type(c_ptr) :: seaice = c_null_ptr, sss = c_null_ptr, fresh = c_null_ptr
      type(c_ptr) :: restore = c_null_ptr
   end type licom_fields

   type, bind(C), public :: licom_diag
      type(c_ptr) :: ax = c_null_ptr, ay = c_null_ptr, az = c_null_ptr, dz = c_null_ptr
      type(c_ptr) :: tend = c_null_ptr, dt_diff = c_null_ptr, dt_conv = c_null_ptr
      type(c_ptr) :: penetrate = c_null_ptr, net = c_null_ptr, fw_norm2 = c_null_ptr
   end type licom_diag

   public :: licom_tracer_create, licom_tracer_set_stream, licom_tracer_step, &
             licom_convadj, licom_icesnow, licom_tracer_set_icesnow, licom_tracer_sync, &
             licom_tracer_destroy

   interface
      integer(c_int) function licom_tracer_create(grid, params, ctx) bind(C, name="licom_tracer_create")
         import :: c_int, c_ptr, licom_grid, licom_params
         type(licom_grid), intent(in) :: grid
         type(licom_params), intent(in) :: params
         type(c_ptr), intent(out) :: ctx
      end function licom_tracer_create

      integer(c_int) function licom_tracer_set_stream(ctx, stream) bind(C, name="licom_tracer_set_stream")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, stream
      end function licom_tracer_set_stream

      integer(c_int) function licom_tracer_step(ctx, fields, diag, flags) bind(C, name="licom_tracer_step")
         import :: c_int, c_ptr, licom_fields, licom_diag
         type(c_ptr), value :: ctx
         type(licom_fields), intent(in) :: fields
         type(licom_diag), intent(in) :: diag
         integer(c_int), value :: flags
      end function licom_tracer_step

      integer(c_int) function licom_convadj(ctx, fields, diag) bind(C, name="licom_convadj")
         import :: c_int, c_ptr, licom_fields, licom_diag
         type(c_ptr), value :: ctx
         type(licom_fields), intent(in) :: fields
         type(licom_diag), intent(in) :: diag
      end function licom_convadj

      ! icesnow_ (icesnow.cpp:70-85); at, licomqice are device addresses
      integer(c_int) function licom_icesnow(ctx, at, licomqice, reset_qice) bind(C, name="licom_icesnow")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, at, licomqice
         integer(c_int), value :: reset_qice
      end function licom_icesnow

      integer(c_int) function licom_tracer_set_icesnow(ctx, licomqice) bind(C, name="licom_tracer_set_icesnow")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, licomqice
      end function licom_tracer_set_icesnow

      integer(c_int) function licom_tracer_sync(ctx) bind(C, name="licom_tracer_sync")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function licom_tracer_sync

      subroutine licom_tracer_destroy(ctx) bind(C, name="licom_tracer_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine licom_tracer_destroy
   end interface
end module licom_tracer_mod
