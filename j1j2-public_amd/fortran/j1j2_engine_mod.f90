! j1j2_engine_mod.f90 -- Fortran binding of the engine's C-ABI
! (include/j1j2_engine.h) through iso_c_binding.
!
! The reference program (sim/J1J2_6tv.f90) calls its hot path as internal
! subroutines; a Fortran host replaces those calls with the functions below:
!   Initialize_pt_state      (J1J2_6tv.f90:1218-1247) -> j1j2_set_spins
!   Pt_equilibration_loop    (J1J2_6tv.f90:1250-1279) -> j1j2_run(h, J1J2_PHASE_EQUIL, -1)
!   Pt_measurement_loop      (J1J2_6tv.f90:1341-1685) -> j1j2_run(h, J1J2_PHASE_MEAS, -1)
!   Collect_data packing     (J1J2_6tv.f90:2937-3000) -> j1j2_fetch
!   Setup_phi_table          (J1J2_6tv.f90:590-633)   -> j1j2_make_disorder + j1j2_upload_disorder
!   Setup_temperature_ladder (J1J2_6tv.f90:928-962)   -> j1j2_ladder
!   Validate_run_options     (J1J2_6tv.f90:841-925)   -> status of j1j2_create_batch
! Every function returns 0 or a negative status; j1j2_check turns a failure
! into the reference's behaviour on invalid options (message, stop 1).
module j1j2_engine_mod
   use, intrinsic :: iso_c_binding
   implicit none
   private :: c_to_f_string

   integer(c_int32_t), parameter :: J1J2_SAMPLER_PT = 0, J1J2_SAMPLER_LOCAL = 1
   integer(c_int32_t), parameter :: J1J2_FP64_PLAIN = 0, J1J2_FP32_BATCHED = 1
   integer(c_int32_t), parameter :: J1J2_LADDER_LINEAR = 0, J1J2_LADDER_SPLIT = 1, J1J2_LADDER_LOCAL = 2
   integer(c_int32_t), parameter :: J1J2_PHASE_EQUIL = 0, J1J2_PHASE_MEAS = 1
   integer(c_int32_t), parameter :: J1J2_FLAG_WRITELOCNEM = 1, J1J2_FLAG_WRITEHIST = 2
   integer(c_int32_t), parameter :: J1J2_FLAG_WRITENUMFLIPS = 4, J1J2_FLAG_SHADOW_FP64 = 8
   integer(c_int32_t), parameter :: J1J2_FLAG_NO_GRAPH = 16, J1J2_FLAG_FORCE_GLOBAL = 32
   integer(c_int32_t), parameter :: J1J2_START_HOT = 0, J1J2_START_FER = 1, J1J2_START_STR = 2
   integer, parameter :: J1J2_BASEFIELDS = 49

   ! struct j1j2_params: 14 int32, 9 double, 1 uint64 (C layout, no padding)
   type, bind(C) :: j1j2_params
      integer(c_int32_t) :: L, H, nreplica, ptswapever, ptmeasindex, neq, nmess, nconf, iconf0, &
                            bins, sampler, precision, ladder, flags
      real(c_double)     :: J1, deltaJ1, J2, J3, tmin, tmax, dt, tladdermid, tladderfrac
      integer(c_int64_t) :: seed
   end type j1j2_params

   interface
      integer(c_int) function j1j2_ladder(p, temps) bind(C, name='j1j2_ladder')
         import :: c_int, c_double, j1j2_params
         type(j1j2_params), intent(in) :: p
         real(c_double), intent(out) :: temps(*)
      end function

      integer(c_int) function j1j2_make_disorder(L, H, xi, seed, map_to_uniform, phi, out) &
         bind(C, name='j1j2_make_disorder')
         import :: c_int, c_int32_t, c_double
         integer(c_int32_t), value :: L, H, seed, map_to_uniform
         real(c_double), value :: xi, phi
         real(c_double), intent(out) :: out(*)
      end function

      integer(c_int) function j1j2_make_disorder_batch(L, H, xi, seed0, nconf, map_to_uniform, phi, nthreads, &
                                                       out) bind(C, name='j1j2_make_disorder_batch')
         import :: c_int, c_int32_t, c_double
         integer(c_int32_t), value :: L, H, seed0, nconf, map_to_uniform, nthreads
         real(c_double), value :: xi, phi
         real(c_double), intent(out) :: out(*)
      end function

      integer(c_int) function j1j2_initial_spins(L, H, start, straxis, seed, skip, spins) &
         bind(C, name='j1j2_initial_spins')
         import :: c_int, c_int32_t, c_int64_t, c_int8_t
         integer(c_int32_t), value :: L, H, start, straxis, seed
         integer(c_int64_t), value :: skip
         integer(c_int8_t), intent(out) :: spins(*)
      end function

      integer(c_int) function j1j2_create(p, device, handle) bind(C, name='j1j2_create')
         import :: c_int, c_int32_t, c_ptr, j1j2_params
         type(j1j2_params), intent(in) :: p
         integer(c_int32_t), value :: device
         type(c_ptr), intent(out) :: handle
      end function

      integer(c_int) function j1j2_create_batch(groups, ngroups, device, handle) &
         bind(C, name='j1j2_create_batch')
         import :: c_int, c_int32_t, c_ptr, j1j2_params
         type(j1j2_params), intent(in) :: groups(*)
         integer(c_int32_t), value :: ngroups, device
         type(c_ptr), intent(out) :: handle
      end function

      subroutine j1j2_destroy(handle) bind(C, name='j1j2_destroy')
         import :: c_ptr
         type(c_ptr), value :: handle
      end subroutine

      integer(c_int) function j1j2_last_error(buf, len) bind(C, name='j1j2_last_error')
         import :: c_int, c_int32_t, c_char
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int32_t), value :: len
      end function

      integer(c_int) function j1j2_num_configs(handle, nconf_total) bind(C, name='j1j2_num_configs')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), intent(out) :: nconf_total
      end function

      integer(c_int) function j1j2_config_shape(handle, conf, L, H, nreplica) &
         bind(C, name='j1j2_config_shape')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         integer(c_int32_t), intent(out) :: L, H, nreplica
      end function

      integer(c_int) function j1j2_upload_disorder(handle, conf, p, n) bind(C, name='j1j2_upload_disorder')
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         real(c_double), intent(in) :: p(*)
         integer(c_int64_t), value :: n
      end function

      integer(c_int) function j1j2_set_spins(handle, conf, spins, nreplica) bind(C, name='j1j2_set_spins')
         import :: c_int, c_int32_t, c_int8_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf, nreplica
         integer(c_int8_t), intent(in) :: spins(*)
      end function

      integer(c_int) function j1j2_set_temperatures(handle, conf, temps, nreplica) &
         bind(C, name='j1j2_set_temperatures')
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf, nreplica
         real(c_double), intent(in) :: temps(*)
      end function

      integer(c_int) function j1j2_run(handle, phase, ncycles) bind(C, name='j1j2_run')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: phase, ncycles
      end function

      integer(c_int) function j1j2_run_cycles(handle, phase, ncycles, begin) bind(C, name='j1j2_run_cycles')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: phase, ncycles, begin
      end function

      integer(c_int) function j1j2_fetch(handle, conf, basefields, hist, hist_min, hist_bin, swap_att, &
                                         swap_acc, visitboth, locspin, locnem, locnem_samples) &
         bind(C, name='j1j2_fetch')
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         real(c_double), intent(out) :: basefields(*), hist(*), hist_min(*), hist_bin(*)
         integer(c_int32_t), intent(out) :: swap_att(*), swap_acc(*), visitboth
         real(c_double), intent(out) :: locspin(*), locnem(*)
         integer(c_int32_t), intent(out) :: locnem_samples
      end function

      integer(c_int) function j1j2_get_spins(handle, conf, spins, nreplica) bind(C, name='j1j2_get_spins')
         import :: c_int, c_int32_t, c_int8_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf, nreplica
         integer(c_int8_t), intent(out) :: spins(*)
      end function

      integer(c_int) function j1j2_get_maps(handle, conf, rep_at_temp, temp_of_rep) bind(C, name='j1j2_get_maps')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         integer(c_int32_t), intent(out) :: rep_at_temp(*), temp_of_rep(*)
      end function

      integer(c_int) function j1j2_get_observables(handle, conf, out) bind(C, name='j1j2_get_observables')
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         real(c_double), intent(out) :: out(*)
      end function

      integer(c_int) function j1j2_get_sums(handle, conf, sums, nmeasure, best, final_en) &
         bind(C, name='j1j2_get_sums')
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         real(c_double), intent(out) :: sums(*), best(*), final_en(*)
         integer(c_int64_t), intent(out) :: nmeasure
      end function

      integer(c_int) function j1j2_get_series(handle, conf, series, cap, len) bind(C, name='j1j2_get_series')
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         real(c_double), intent(out) :: series(*)
         integer(c_int64_t), value :: cap
         integer(c_int64_t), intent(out) :: len
      end function

      integer(c_int) function j1j2_get_flips(handle, conf, flips) bind(C, name='j1j2_get_flips')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         integer(c_int32_t), intent(out) :: flips(*)
      end function

      integer(c_int) function j1j2_get_mismatch(handle, conf, proposals, mismatches) &
         bind(C, name='j1j2_get_mismatch')
         import :: c_int, c_int32_t, c_int64_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: conf
         integer(c_int64_t), intent(out) :: proposals, mismatches
      end function

      integer(c_int) function j1j2_set_timing(handle, enable) bind(C, name='j1j2_set_timing')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: enable
      end function

      integer(c_int) function j1j2_get_timing(handle, kind, total_ms, launches) bind(C, name='j1j2_get_timing')
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), value :: kind
         real(c_double), intent(out) :: total_ms
         integer(c_int64_t), intent(out) :: launches
      end function

      integer(c_int) function j1j2_sweep_path(handle, path, bytes_per_update) bind(C, name='j1j2_sweep_path')
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: handle
         integer(c_int32_t), intent(out) :: path
         real(c_double), intent(out) :: bytes_per_update
      end function

      integer(c_int) function j1j2_launch_plan(handle, buf, len) bind(C, name='j1j2_launch_plan')
         import :: c_int, c_int32_t, c_char, c_ptr
         type(c_ptr), value :: handle
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int32_t), value :: len
      end function

      ! no counterpart in the reference: for callers that cannot allocate
      integer(c_int) function j1j2_run_cycles_noalloc(handle, phase, ncycles, begin, stream) &
         bind(C, name='j1j2_run_cycles_noalloc')
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: handle, stream
         integer(c_int32_t), value :: phase, ncycles, begin
      end function

      integer(c_int) function j1j2_synchronize(handle) bind(C, name='j1j2_synchronize')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function
   end interface

contains

   ! Message of the last failed call, as a Fortran string.
   function j1j2_error_message() result(msg)
      character(len=:), allocatable :: msg
      character(kind=c_char) :: buf(512)
      integer(c_int) :: rc
      rc = j1j2_last_error(buf, 512_c_int32_t)
      msg = c_to_f_string(buf)
   end function

   ! A non-zero status ends the run as the reference ends on an invalid
   ! option (J1J2_6tv.f90:841-925): the message, then stop 1.
   subroutine j1j2_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= 0) then
         print *, 'ERROR: ', trim(where), ': ', j1j2_error_message()
         stop 1
      end if
   end subroutine

   function c_to_f_string(buf) result(s)
      character(kind=c_char), intent(in) :: buf(:)
      character(len=:), allocatable :: s
      integer :: n, i
      n = 0
      do while (n < size(buf))
         if (buf(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate (character(len=n) :: s)
      do i = 1, n
         s(i:i) = buf(i)
      end do
   end function
end module j1j2_engine_mod
