! j1j2_host.f90 -- Fortran host of the engine: the driver of program J1J2_6
! (J1J2_6tv.f90:268-405) with its hot path on the GPU.
!
! Run parameters are the reference's parameter block (J1J2_6tv.f90:43-96),
! read from a namelist file (argument 1, default j1j2.nml) instead of being
! compiled in, plus three engine choices: PRECISION ('FP32' batched form or
! 'FP64' plain form), DEVSEED (Philox key of the device RNG, -1 = IRINIT),
! DEVICE, and BATCH (disorder configs per engine launch, 0 = all).
!
! Per disorder config iconf the host does what the reference does before its
! hot path -- the correlated field for seed IRINIT+iconf-1 (Setup_phi_table,
! J1J2_6tv.f90:590-633) and the starts of Initialize_pt_state (J1J2_6tv.f90:
! 1218-1247) or of the LOCAL protocols (J1J2_6tv.f90:316-322, 1098-1106) --
! then one engine runs every config of a batch at once.  The results are
! collected as Collect_data (J1J2_6tv.f90:2917-3207, serial branch) and written
! in the layout of Write_files (J1J2_6tv.f90:3210-3575): cfen, cfma, cfnm,
! cfst, cfhs, cfxy, cfnd, cfnu, cfpt, cfph, plus phi_sites.dat of the last
! config and the locnem files of config 1 (Local_nematic_OP,
! J1J2_6tv.f90:2845-2884).  Files are rewritten after every batch with the
! number of configs completed so far, as the reference rewrites them after
! every config.  progress.txt and the progress lines follow Write_progress
! (J1J2_6tv.f90:499-529; a PT phase runs in chunks of a tenth, as the
! reference reports), and the timing lines Write_timing (J1J2_6tv.f90:475-496)
! per batch: iconf is the batch's first config and a trailing field names
! the batch's configs.
!
! Not supported (rejected with a message and stop 1): dilution (PP, PPh,
! PPv > 0), UPDATEORDER = 'RND', PTPARALLELMODE = 'GROUPED'.  WRITEBINDTIME
! trace files (cfbt) are not written.
program j1j2_host
   use, intrinsic :: iso_c_binding
   use j1j2_engine_mod
   implicit none

   integer, parameter :: r8b = c_double, i4b = c_int32_t, i1b = c_int8_t

   ! Position (1-based) of each quantity in the 49-field block of
   ! Collect_data's packing (J1J2_6tv.f90:2937-2985).
   integer, parameter :: F_MAG = 1, F_MAG2 = 2, F_MAG4 = 3, F_SUSC = 4, F_2SUSC = 5, F_UL = 6, &
      F_2UL = 7, F_2MAG = 8, F_LOGMAG = 9, F_EN = 10, F_2EN = 11, F_SPH = 12, F_2SPH = 13, &
      F_ABSSTR = 14, F_STRX = 15, F_STRY = 16, F_STR2 = 17, F_STR4 = 18, F_SUSCSTR = 19, &
      F_2SUSCSTR = 20, F_US = 21, F_2US = 22, F_NEM = 23, F_2NEM = 24, F_2STR = 25, F_STRX2 = 26, &
      F_STRX4 = 27, F_USX = 28, F_2USX = 29, F_STRY2 = 30, F_STRY4 = 31, F_USY = 32, F_2USY = 33, &
      F_ABSSTRD = 34, F_STRDIFF = 35, F_BESTEN = 36, F_FINALEN = 37, F_NBAR = 38, F_2NBAR = 39, &
      F_NBAR2 = 40, F_NBAR4 = 41, F_UBN = 42, F_2UBN = 43, F_CNBAR = 44, F_CNBAR2 = 45, &
      F_CNBAR4 = 46, F_CNBIND = 47, F_NBARABS = 48, F_CNBARABS = 49

   ! ---- run parameters (reference defaults)
   integer(i4b) :: L = 10, H = 10
   real(r8b) :: J1 = 1.0D0, deltaJ1 = 0.0D0, J2 = -1.0D0, J3 = 1.0D0
   real(r8b) :: phi = 2.0D0, corr_length = 3.0D0
   real(r8b) :: TMAX = 14.0D0, TMIN = 8.0D0, DT = 2.0D0
   real(r8b) :: PP = 0.0D0, PPh = 0.0D0, PPv = 0.0D0
   character(len=3) :: IMPMODE = 'RAN', START = 'STR', UPDATEORDER = 'SEQ'
   character(len=1) :: STRAXIS = 'X'
   character(len=5) :: TEMPPROTOCOL = 'CARRY', SAMPLERMODE = 'LOCAL'
   character(len=6) :: PTSTARTMODE = 'COMMON', PTLADDERMODE = 'LINEAR'
   character(len=7) :: PTPARALLELMODE = 'SERIAL'
   real(r8b) :: TLADDERMID = 9.5D0, TLADDERFRAC = 0.7D0
   integer(i4b) :: NEQ = 1000000, NMESS = 1000000, NCONF = 100, PTSWAPEVER = 5
   logical :: WRITELOCNEM = .true., WRITEHIST = .true., WRITESTRIP = .true., WRITENUMFLIPS = .false.
   logical :: WRITEBINDTIME = .true., WRITEPHITABLE = .true., MAPTOUNIFORM = .true.
   logical :: VALIDATE_INCREMENTAL = .false.
   integer(i4b) :: BINS = 100, NTRACECONF = 5, IRINIT = 1, NREPLICA = 1, PTMEASINDEX = 1
   character(len=4) :: PRECISION = 'FP32'
   integer(c_int64_t) :: DEVSEED = -1
   integer(i4b) :: DEVICE = 0, BATCH = 0

   namelist /j1j2/ L, H, J1, deltaJ1, J2, J3, phi, corr_length, TMAX, TMIN, DT, PP, PPh, PPv, &
      IMPMODE, START, STRAXIS, UPDATEORDER, TEMPPROTOCOL, SAMPLERMODE, PTSTARTMODE, PTLADDERMODE, &
      PTPARALLELMODE, TLADDERMID, TLADDERFRAC, NEQ, NMESS, NCONF, PTSWAPEVER, WRITELOCNEM, WRITEHIST, &
      WRITESTRIP, WRITENUMFLIPS, WRITEBINDTIME, WRITEPHITABLE, MAPTOUNIFORM, VALIDATE_INCREMENTAL, &
      BINS, NTRACECONF, IRINIT, NREPLICA, PTMEASINDEX, PRECISION, DEVSEED, DEVICE, BATCH

   ! ---- run state
   integer(i4b) :: NTEMP, L3
   type(j1j2_params) :: base
   real(r8b), allocatable :: temp_ladder(:)
   real(r8b), allocatable :: cfgfields(:, :, :)        ! (NTEMP, 49, NCONF)
   real(r8b), allocatable :: globalhist(:, :)          ! (NTEMP, BINS)
   real(r8b), allocatable :: hist_min(:), hist_bin(:)  ! of the last config processed
   integer(i4b), allocatable :: totalnumflips(:, :, :) ! (0:L3-1, NTEMP, NCONF)
   real(r8b), allocatable :: jointp(:, :)              ! (0:L3-1, NCONF)
   integer(c_int64_t), allocatable :: sum_att(:), sum_acc(:)
   integer(c_int64_t) :: sum_visitboth = 0
   integer(i4b) :: first, nbatch, last
   integer(c_int64_t) :: tc0, tc1, tc2, trate

   call read_options()
   call validate_options()

   L3 = L*L*H
   if (SAMPLERMODE == 'PT') then
      NTEMP = NREPLICA
   else
      NTEMP = 1 + nint((TMAX - TMIN)/DT)
   end if
   call setup_base()
   allocate (temp_ladder(NTEMP))
   call setup_ladder()

   allocate (cfgfields(NTEMP, J1J2_BASEFIELDS, NCONF), globalhist(NTEMP, BINS))
   allocate (hist_min(NTEMP), hist_bin(NTEMP), sum_att(max(1, NTEMP - 1)), sum_acc(max(1, NTEMP - 1)))
   cfgfields = 0.D0; globalhist = 0.D0; hist_min = 0.D0; hist_bin = 0.D0; sum_att = 0; sum_acc = 0
   if (WRITENUMFLIPS) then
      allocate (totalnumflips(0:L3 - 1, NTEMP, NCONF))
      totalnumflips = 0
   end if
   if (WRITEPHITABLE) then
      allocate (jointp(0:L3 - 1, NCONF))
      jointp = 0.D0
   end if

   nbatch = BATCH
   if (nbatch <= 0) nbatch = NCONF
   do first = 1, NCONF, nbatch
      last = min(NCONF, first + nbatch - 1)
      call system_clock(tc0, trate)
      call write_progress('CONFIG_START', first, 0, 0)
      if (SAMPLERMODE == 'PT') then
         call run_pt_batch(first, last - first + 1)
      else
         call run_local_batch(first, last - first + 1)
      end if
      ! Collect_data + Write_files of the reference (J1J2_6tv.f90:373-384)
      call system_clock(tc1)
      call write_files(last)
      call system_clock(tc2)
      call print_timing('COLLECT_DATA', first, last, tc1, tc2, trate)
      call print_timing('CONFIG_TOTAL', first, last, tc0, tc2, trate)
   end do

contains

   ! ------------------------------------------------------------ options
   subroutine read_options()
      character(len=512) :: path
      integer :: u, ios
      if (command_argument_count() >= 1) then
         call get_command_argument(1, path)
      else
         path = 'j1j2.nml'
      end if
      open (newunit=u, file=trim(path), status='old', action='read', iostat=ios)
      if (ios /= 0) then
         print *, 'ERROR: cannot open namelist file ', trim(path)
         stop 1
      end if
      read (u, nml=j1j2, iostat=ios)
      if (ios /= 0) then
         print *, 'ERROR: cannot read namelist j1j2 from ', trim(path)
         stop 1
      end if
      close (u)
      if (DEVSEED < 0) DEVSEED = IRINIT
   end subroutine

   subroutine fail(msg)
      character(len=*), intent(in) :: msg
      print *, 'ERROR: ', msg
      stop 1
   end subroutine

   ! The string-valued checks of Validate_run_options (J1J2_6tv.f90:841-925)
   ! live here; the numeric ones are the engine's (status of j1j2_create).
   subroutine validate_options()
      if (SAMPLERMODE /= 'LOCAL' .and. SAMPLERMODE /= 'PT') &
         call fail('SAMPLERMODE must be LOCAL or PT. Got '//trim(SAMPLERMODE))
      if (UPDATEORDER /= 'SEQ' .and. UPDATEORDER /= 'RND') &
         call fail('UPDATEORDER must be SEQ or RND. Got '//trim(UPDATEORDER))
      if (TEMPPROTOCOL /= 'CARRY' .and. TEMPPROTOCOL /= 'INDEP') &
         call fail('TEMPPROTOCOL must be CARRY or INDEP. Got '//trim(TEMPPROTOCOL))
      if (PTSTARTMODE /= 'COMMON' .and. PTSTARTMODE /= 'INDEP') &
         call fail('PTSTARTMODE must be COMMON or INDEP. Got '//trim(PTSTARTMODE))
      if (PTLADDERMODE /= 'LINEAR' .and. PTLADDERMODE /= 'SPLIT') &
         call fail('PTLADDERMODE must be LINEAR or SPLIT. Got '//trim(PTLADDERMODE))
      if (PTPARALLELMODE /= 'SERIAL' .and. PTPARALLELMODE /= 'GROUPED') &
         call fail('PTPARALLELMODE must be SERIAL or GROUPED. Got '//trim(PTPARALLELMODE))
      if (TMAX < TMIN) call fail('TMAX must be >= TMIN.')
      if (SAMPLERMODE /= 'PT' .and. DT <= 0.D0) call fail('DT must be positive.')
      if (PTPARALLELMODE == 'GROUPED') &
         call fail('PTPARALLELMODE=GROUPED is not supported: one GPU holds the whole ladder.')
      if (UPDATEORDER == 'RND') &
         call fail('UPDATEORDER=RND is not supported: the engine sweeps in 4-colour order.')
      if (PP > 0.D0 .or. PPh > 0.D0 .or. PPv > 0.D0) &
         call fail('dilution (PP, PPh, PPv > 0) is not supported by the engine.')
      if (PRECISION /= 'FP32' .and. PRECISION /= 'FP64') &
         call fail('PRECISION must be FP32 or FP64. Got '//trim(PRECISION))
      if (NCONF < 1) call fail('NCONF must be positive.')
      if (WRITEBINDTIME) print *, 'note: WRITEBINDTIME trace files (cfbt) are not written'
   end subroutine

   subroutine setup_base()
      base%L = L; base%H = H
      base%nreplica = NTEMP
      base%ptswapever = PTSWAPEVER
      if (SAMPLERMODE /= 'PT') base%ptswapever = 1     ! LOCAL measures after every sweep
      base%ptmeasindex = PTMEASINDEX
      base%neq = NEQ; base%nmess = NMESS
      base%nconf = 1; base%iconf0 = 1
      base%bins = BINS
      if (SAMPLERMODE == 'PT') then
         base%sampler = J1J2_SAMPLER_PT
         if (PTLADDERMODE == 'SPLIT') then
            base%ladder = J1J2_LADDER_SPLIT
         else
            base%ladder = J1J2_LADDER_LINEAR
         end if
      else
         base%sampler = J1J2_SAMPLER_LOCAL
         base%ladder = J1J2_LADDER_LOCAL
      end if
      if (PRECISION == 'FP64') then
         base%precision = J1J2_FP64_PLAIN
      else
         base%precision = J1J2_FP32_BATCHED
      end if
      base%flags = 0
      if (WRITELOCNEM) base%flags = ior(base%flags, J1J2_FLAG_WRITELOCNEM)
      if (WRITEHIST) base%flags = ior(base%flags, J1J2_FLAG_WRITEHIST)
      if (WRITENUMFLIPS) base%flags = ior(base%flags, J1J2_FLAG_WRITENUMFLIPS)
      base%J1 = J1; base%deltaJ1 = deltaJ1; base%J2 = J2; base%J3 = J3
      base%tmin = TMIN; base%tmax = TMAX; base%dt = DT
      base%tladdermid = TLADDERMID; base%tladderfrac = TLADDERFRAC
      base%seed = DEVSEED
   end subroutine

   ! Setup_temperature_ladder (J1J2_6tv.f90:928-962), computed by the library.
   subroutine setup_ladder()
      call j1j2_check(j1j2_ladder(base, temp_ladder), 'temperature ladder')
   end subroutine

   ! Temperature_for_index (J1J2_6tv.f90:987-996): LOCAL runs with a HOT
   ! start walk the ladder downwards.
   real(r8b) function temperature_for_index(it)
      integer, intent(in) :: it
      if (SAMPLERMODE == 'PT') then
         temperature_for_index = temp_ladder(it)
      else if (START == 'HOT') then
         temperature_for_index = temp_ladder(NTEMP - it + 1)
      else
         temperature_for_index = temp_ladder(it)
      end if
   end function

   integer(i4b) function start_code()
      if (START == 'HOT') then
         start_code = J1J2_START_HOT
      else if (START == 'FER') then
         start_code = J1J2_START_FER
      else
         start_code = J1J2_START_STR
      end if
   end function

   integer(i4b) function axis_code()
      if (STRAXIS == 'X') then
         axis_code = 0
      else
         axis_code = 1
      end if
   end function

   ! Setup_phi_table (J1J2_6tv.f90:590-633) for the configs first .. first+n-1
   ! at once (seeds IRINIT + iconf - 1, on all host cores).
   subroutine make_fields(first, n, fields)
      integer, intent(in) :: first, n
      real(r8b), intent(out) :: fields(0:L3 - 1, n)
      integer(i4b) :: map
      map = 0
      if (MAPTOUNIFORM) map = 1
      call j1j2_check(j1j2_make_disorder_batch(L, H, corr_length, IRINIT + first - 1, int(n, i4b), map, phi, &
                                               0_i4b, fields), 'correlated field')
   end subroutine

   subroutine upload_field(eng, k, iconf, field)
      type(c_ptr), intent(in) :: eng
      integer, intent(in) :: k, iconf
      real(r8b), intent(in) :: field(0:L3 - 1)
      call j1j2_check(j1j2_upload_disorder(eng, int(k, i4b), field, int(L3, c_int64_t)), 'upload field')
      if (WRITEPHITABLE) jointp(:, iconf) = field
   end subroutine

   subroutine write_phi_sites(field)
      real(r8b), intent(in) :: field(0:L3 - 1)
      integer :: u, ix, iy, iz
      open (newunit=u, file='phi_sites.dat', status='replace')
      write (u, *) '# ix iy layer isite phi'
      do iz = 0, H - 1
         do iy = 0, L - 1
            do ix = 0, L - 1
               write (u, '(3I8,1X,E20.12)') ix, iy, iz, field(ix + L*iy + L*L*iz)
            end do
         end do
      end do
      close (u)
   end subroutine

   ! The reference's timing line (Write_timing, J1J2_6tv.f90:475-496): the
   ! phase's seconds for the batch of configs a .. b run in one engine, so
   ! iconf is the batch's first config and the configs of the batch follow.
   subroutine print_timing(label, a, b, t0, t1, rate)
      character(len=*), intent(in) :: label
      integer, intent(in) :: a, b
      integer(c_int64_t), intent(in) :: t0, t1, rate
      write (*, '(A,1x,A,1x,A,I0,1x,A,I0,1x,A,F12.3,1x,A,I0,A,I0)') 'timing', trim(label), 'rank', 0, &
         'iconf', a, 'seconds', real(t1 - t0, r8b)/real(rate, r8b), 'configs ', a, '-', b
      flush (6)
   end subroutine

   ! progress.txt and the progress line of the reference (Write_progress,
   ! J1J2_6tv.f90:499-529), for the batch whose first config is iconf.
   subroutine write_progress(phase, iconf, current, total)
      character(len=*), intent(in) :: phase
      integer, intent(in) :: iconf, current, total
      character(len=64) :: job_id
      integer :: u
      job_id = ''
      call get_environment_variable('SLURM_JOB_ID', job_id)
      open (newunit=u, file='progress.txt', status='replace')
      write (u, '(A,1x,A)') 'job_id', trim(job_id)
      write (u, '(A,1x,A)') 'sampler', trim(SAMPLERMODE)
      write (u, '(A,1x,A)') 'phase', trim(phase)
      write (u, '(A,1x,I0)') 'iconf', iconf
      write (u, '(A,1x,I0)') 'nconf', NCONF
      write (u, '(A,1x,I0)') 'ntasks', 1
      write (u, '(A,1x,I0)') 'current', current
      write (u, '(A,1x,I0)') 'total', total
      close (u)
      write (*, '(A,1x,A,1x,A,I0,1x,A,I0,1x,A,I0)') 'progress', trim(phase), 'iconf', iconf, &
         'step', current, 'of', total
      flush (6)
   end subroutine

   ! One PT phase of `total` cycles in chunks of max(1, total/10), with the
   ! reference's progress reports (J1J2_6tv.f90:1267-1276, 1430-1531); the
   ! chunks continue one run (the first begins the phase), so the result is
   ! that of a single j1j2_run.  The first chunk holds at least TUNE_CYCLES
   ! cycles (a whole number of tenths) when the phase has that many, so that
   ! the engine's autotuner, which times its candidates on the first call long
   ! enough for them, gets to run.  A phase of no cycles is one j1j2_run, which
   ! samples the state as the reference's NMESS <= 0 branch does.
   subroutine run_pt_phase(eng, phase, label, iconf, total)
      type(c_ptr), intent(in) :: eng
      integer(c_int), intent(in) :: phase
      character(len=*), intent(in) :: label
      integer, intent(in) :: iconf, total
      integer, parameter :: TUNE_CYCLES = 60
      integer :: every, done, chunk
      every = max(1, total/10)
      call write_progress(label, iconf, 0, total)
      if (total <= 0) then
         call j1j2_check(j1j2_run(eng, phase, 0_i4b), label)
         return
      end if
      done = 0
      chunk = min(every, total)
      if (total >= TUNE_CYCLES .and. chunk < TUNE_CYCLES) chunk = min(total, every*((TUNE_CYCLES + every - 1)/every))
      call j1j2_check(j1j2_run_cycles(eng, phase, int(chunk, i4b), 1_i4b), label)
      done = chunk
      if (done > 0) call write_progress(label, iconf, done, total)
      do while (done < total)
         chunk = min(every, total - done)
         call j1j2_check(j1j2_run_cycles(eng, phase, int(chunk, i4b), 0_i4b), label)
         done = done + chunk
         call write_progress(label, iconf, done, total)
      end do
   end subroutine

   ! ------------------------------------------------------------ PT
   ! PT SERIAL for configs first .. first+n-1 (J1J2_6tv.f90:340-344).
   subroutine run_pt_batch(first, n)
      integer, intent(in) :: first, n
      type(j1j2_params) :: p
      type(c_ptr) :: eng
      real(r8b), allocatable :: fields(:, :), bf(:, :), hist(:, :), hmin(:), hbin(:), locspin(:), locnem(:)
      integer(i1b), allocatable :: spins(:, :)
      integer(i4b), allocatable :: att(:), acc(:), flips(:, :)
      integer(i4b) :: visit, nsamp
      integer :: k, iconf, rep
      integer(c_int64_t) :: t0, t1, t2, rate

      p = base
      p%nconf = n
      p%iconf0 = first
      call j1j2_check(j1j2_create(p, DEVICE, eng), 'engine')
      allocate (fields(0:L3 - 1, n), spins(0:L3 - 1, NTEMP), bf(NTEMP, J1J2_BASEFIELDS), hist(NTEMP, BINS))
      allocate (hmin(NTEMP), hbin(NTEMP), locspin(0:L3 - 1), locnem(0:L3 - 1))
      allocate (att(max(1, NTEMP - 1)), acc(max(1, NTEMP - 1)), flips(0:L3 - 1, NTEMP))

      call make_fields(first, n, fields)
      do k = 0, n - 1
         iconf = first + k
         print *, 'dis. conf.', iconf
         call upload_field(eng, k, iconf, fields(:, k + 1))
         ! Initialize_pt_state: replica rep starts from KISS seeded with
         ! IRINIT + (iconf-1)*NREPLICA + rep.
         do rep = 1, NTEMP
            call j1j2_check(j1j2_initial_spins(L, H, start_code(), axis_code(), &
                            IRINIT + (iconf - 1)*NTEMP + rep, 0_c_int64_t, spins(:, rep)), 'initial spins')
         end do
         call j1j2_check(j1j2_set_spins(eng, int(k, i4b), spins, NTEMP), 'set spins')
      end do
      call write_phi_sites(fields(:, n))

      call system_clock(t0, rate)
      call run_pt_phase(eng, J1J2_PHASE_EQUIL, 'PT_EQUIL', first, NEQ)
      call system_clock(t1)
      call print_timing('PT_EQUIL', first, first + n - 1, t0, t1, rate)
      call run_pt_phase(eng, J1J2_PHASE_MEAS, 'PT_MEAS', first, NMESS)
      call system_clock(t2)
      call print_timing('PT_MEAS', first, first + n - 1, t1, t2, rate)

      do k = 0, n - 1
         iconf = first + k
         call j1j2_check(j1j2_fetch(eng, int(k, i4b), bf, hist, hmin, hbin, att, acc, visit, locspin, locnem, &
                                    nsamp), 'fetch')
         cfgfields(:, :, iconf) = bf
         globalhist = globalhist + hist
         hist_min = hmin
         hist_bin = hbin
         if (NTEMP > 1) then
            sum_att(1:NTEMP - 1) = sum_att(1:NTEMP - 1) + att(1:NTEMP - 1)
            sum_acc(1:NTEMP - 1) = sum_acc(1:NTEMP - 1) + acc(1:NTEMP - 1)
         end if
         sum_visitboth = sum_visitboth + visit
         if (WRITENUMFLIPS) then
            call j1j2_check(j1j2_get_flips(eng, int(k, i4b), flips), 'flip counts')
            totalnumflips(:, :, iconf) = flips
         end if
         ! Snapshot_pt_target_replica + Local_nematic_OP for config 1
         if (WRITELOCNEM .and. iconf == 1) &
            call write_locnem(PTMEASINDEX, temp_ladder(PTMEASINDEX), bf(PTMEASINDEX, F_STRX), &
                              bf(PTMEASINDEX, F_STRY), locspin, locnem, nsamp)
      end do
      call j1j2_destroy(eng)
   end subroutine

   ! ------------------------------------------------------------ LOCAL
   ! LOCAL sampler for configs first .. first+n-1, one chain per config,
   ! temperatures in the order of Temperature_for_index (J1J2_6tv.f90:346-361).
   subroutine run_local_batch(first, n)
      integer, intent(in) :: first, n
      type(j1j2_params) :: p
      type(c_ptr) :: eng
      real(r8b), allocatable :: fields(:, :), bf(:, :), hist(:, :), locspin(:), locnem(:)
      integer(i1b), allocatable :: spins(:)
      integer(i4b), allocatable :: flips(:), prev(:, :)
      real(r8b) :: hmin(1), hbin(1), t(1)
      integer(i4b) :: att(1), acc(1), visit, nsamp
      integer :: k, iconf, itemp
      integer(c_int64_t) :: skip, t0, t1, t2, rate

      p = base
      p%nconf = n
      p%iconf0 = first
      p%nreplica = 1
      p%ptswapever = 1
      p%ptmeasindex = 1
      p%tmin = temperature_for_index(1)
      p%tmax = p%tmin
      call j1j2_check(j1j2_create(p, DEVICE, eng), 'engine')
      allocate (fields(0:L3 - 1, n), spins(0:L3 - 1), bf(1, J1J2_BASEFIELDS), hist(1, BINS))
      allocate (locspin(0:L3 - 1), locnem(0:L3 - 1), flips(0:L3 - 1), prev(0:L3 - 1, 0:n - 1))
      prev = 0

      ! KISS draws made by Setup_vacancies before the carried start
      ! (J1J2_6tv.f90:1116-1175): occupancy (IMPMODE RAN only), then the
      ! horizontal and vertical bond masks, one draw per site each.
      skip = 2_c_int64_t*L3
      if (IMPMODE == 'RAN') skip = 3_c_int64_t*L3
      call make_fields(first, n, fields)
      do k = 0, n - 1
         iconf = first + k
         print *, 'dis. conf.', iconf
         call upload_field(eng, k, iconf, fields(:, k + 1))
         if (TEMPPROTOCOL == 'CARRY') then
            call j1j2_check(j1j2_initial_spins(L, H, start_code(), axis_code(), IRINIT + iconf - 1, skip, &
                                               spins), 'initial spins')
            call j1j2_check(j1j2_set_spins(eng, int(k, i4b), spins, 1_i4b), 'set spins')
         end if
      end do
      call write_phi_sites(fields(:, n))

      do itemp = 1, NTEMP
         t(1) = temperature_for_index(itemp)
         print *, 'Running T= ', t(1), 'itemp= ', itemp
         do k = 0, n - 1
            iconf = first + k
            if (TEMPPROTOCOL == 'INDEP') then
               ! Prepare_temperature_state: fresh start from a per-temperature seed
               call j1j2_check(j1j2_initial_spins(L, H, start_code(), axis_code(), &
                               IRINIT + (iconf - 1)*NTEMP + itemp, 0_c_int64_t, spins), 'initial spins')
               call j1j2_check(j1j2_set_spins(eng, int(k, i4b), spins, 1_i4b), 'set spins')
               prev(:, k) = 0
            end if
            call j1j2_check(j1j2_set_temperatures(eng, int(k, i4b), t, 1_i4b), 'set temperature')
         end do
         call system_clock(t0, rate)
         call j1j2_check(j1j2_run(eng, J1J2_PHASE_EQUIL, -1_i4b), 'equilibration')
         call system_clock(t1)
         call j1j2_check(j1j2_run(eng, J1J2_PHASE_MEAS, -1_i4b), 'measurement')
         call system_clock(t2)
         call print_timing('EQUIL', first, first + n - 1, t0, t1, rate)
         call print_timing('MEAS', first, first + n - 1, t1, t2, rate)
         do k = 0, n - 1
            iconf = first + k
            call j1j2_check(j1j2_fetch(eng, int(k, i4b), bf, hist, hmin, hbin, att, acc, visit, locspin, &
                                       locnem, nsamp), 'fetch')
            cfgfields(itemp, :, iconf) = bf(1, :)
            globalhist(itemp, :) = globalhist(itemp, :) + hist(1, :)
            hist_min(itemp) = hmin(1)
            hist_bin(itemp) = hbin(1)
            if (WRITENUMFLIPS) then
               ! the engine counts from the last start; the reference per temperature
               call j1j2_check(j1j2_get_flips(eng, int(k, i4b), flips), 'flip counts')
               totalnumflips(:, itemp, iconf) = flips - prev(:, k)
               prev(:, k) = flips
            end if
            if (WRITELOCNEM .and. iconf == 1) &
               call write_locnem(itemp, t(1), bf(1, F_STRX), bf(1, F_STRY), locspin, locnem, nsamp)
         end do
      end do
      call j1j2_destroy(eng)
   end subroutine

   ! Local_nematic_OP (J1J2_6tv.f90:2845-2884): site averages over the
   ! measurement of config 1 at one temperature.
   subroutine write_locnem(itemp, T, strx, stry, locspin, locnem, nsamp)
      integer, intent(in) :: itemp
      real(r8b), intent(in) :: T, strx, stry, locspin(0:L3 - 1), locnem(0:L3 - 1)
      integer(i4b), intent(in) :: nsamp
      character(len=20) :: fname
      integer :: u, ix, iy, iz, isite
      real(r8b) :: norm
      fname = 'locnem0000.dat'
      write (fname(7:10), '(I4.4)') itemp
      norm = real(max(1, nsamp), r8b)
      open (newunit=u, file=trim(fname), status='replace')
      write (u, '(A,E24.16)') '# Temperature = ', T
      write (u, '(A,I0)') '# disorder configuration = ', 1
      write (u, '(A,2(E24.16,1x))') '# stripe magnitudes = ', strx, stry
      write (u, '(A,E24.16)') '# expected mean locnem = ', 4.D0*(stry**2 - strx**2)
      write (u, '(A,E24.16)') '# measured mean locnem = ', sum(locnem)/(norm*real(max(1, L3), r8b))
      write (u, '(A)') '# ix iy iz isite occu <s(isite)> <locnem(isite)>'
      do iz = 0, H - 1
         do iy = 0, L - 1
            do ix = 0, L - 1
               isite = ix + L*iy + L*L*iz
               write (u, '(4(I6,1x),I3,1x,2(E18.10,1x))') ix, iy, iz, isite, 1, locspin(isite)/norm, &
                  locnem(isite)/norm
            end do
         end do
      end do
      close (u)
   end subroutine

   ! ------------------------------------------------------------ output
   ! The run description that heads the averaged files.  kind: 1 = cfen,
   ! 2 = cfma/cfnm/cfst, 3 = cfhs (their headers differ in three lines).
   subroutine write_header(u, kind, finconf)
      integer, intent(in) :: u, kind, finconf
      write (u, *) 'program J1J2_6          '
      if (kind == 2) then
         write (u, *) 'spatial system size     ', L
      else
         write (u, *) 'linear system size      ', L
      end if
      write (u, *) 'J1                      ', J1
      write (u, *) 'deltaJ1                 ', deltaJ1
      write (u, *) 'J2                      ', J2
      write (u, *) 'J3                      ', J3
      if (kind == 3) then
         write (u, *) 'phi                     ', phi
      else
         write (u, *) 'phi, xi                     ', phi, corr_length
      end if
      write (u, *) 'impurity conc.          ', PP
      write (u, *) 'impurity mode           ', IMPMODE
      write (u, *) 'start                   ', START
      write (u, *) 'stripe axis             ', STRAXIS
      write (u, *) 'sampler mode            ', SAMPLERMODE
      write (u, *) 'update order            ', UPDATEORDER
      write (u, *) 'temperature protocol    ', TEMPPROTOCOL
      write (u, *) 'PT start mode           ', PTSTARTMODE
      write (u, *) 'PT ladder mode          ', PTLADDERMODE
      write (u, *) 'PT replicas             ', NREPLICA
      write (u, *) 'PT swap every           ', PTSWAPEVER
      write (u, *) 'PT measure index        ', PTMEASINDEX
      write (u, *) 'disorder configurations ', NCONF
      write (u, *) 'equilibration steps     ', NEQ
      write (u, *) 'measurement steps       ', NMESS
      write (u, *) 'RNG seed                ', IRINIT
      if (PTLADDERMODE == 'SPLIT') then
         write (u, *) 'T ladder midpoint       ', TLADDERMID
         write (u, *) 'T ladder low-T fraction ', TLADDERFRAC
      end if
      if (kind == 2) then
         write (u, *) '------------------------'
      else
         write (u, *) 'energy definition       ', 'exchange + phi'
         write (u, *) '-------------------------'
      end if
      write (u, *) 'disorder configurations processed ', finconf
   end subroutine

   ! Mean over the first n configs of field f at slot it (the sum runs in
   ! config order, as Collect_data accumulates it).  The reference never
   ! clears conf2mag between configs (it is missing from
   ! Reset_current_config_averages, J1J2_6tv.f90:771-838), so what it adds
   ! for config c is the running total of |m|^2 over configs 1..c; the
   ! [<mag>^2] column of cfma is reproduced with that running total.
   real(r8b) function cmean(f, it, n)
      integer, intent(in) :: f, it, n
      real(r8b) :: s, carried
      integer :: c
      s = 0.D0
      carried = 0.D0
      do c = 1, n
         if (f == F_2MAG) then
            carried = carried + cfgfields(it, f, c)
            s = s + carried
         else
            s = s + cfgfields(it, f, c)
         end if
      end do
      cmean = s/n
   end function

   ! Disorder error of a mean from the mean of its square: sqrt(<x^2> - <x>^2)/sqrt(n).
   real(r8b) function cerr(f2, f, it, n)
      integer, intent(in) :: f2, f, it, n
      cerr = (sqrt(cmean(f2, it, n) - cmean(f, it, n)**2))/sqrt(1.D0*n)
   end function

   ! 0.5*(1 - cos(4 theta)) of the stripe vector (strx, stry) (J1J2_6tv.f90:1051-1063).
   real(r8b) function coex_angle(strx, stry)
      real(r8b), intent(in) :: strx, stry
      real(r8b) :: d
      d = strx**2 + stry**2
      if (d > 0.D0) then
         coex_angle = 4.D0*strx**2*stry**2/(d**2)
      else
         coex_angle = 0.D0
      end if
   end function

   subroutine write_files(finconf)
      integer, intent(in) :: finconf
      character(len=8) :: tag
      integer :: u, it, id, ibin, isite
      real(r8b) :: T, a, s, s2, q, q2, m, m2, m4, ubn, e1, e2, e3, mabs
      real(r8b), allocatable :: v(:)

      write (tag, '(I4.4,I4.4)') L, H
      allocate (v(finconf))

      ! ---- cfen: energy and specific heat
      open (newunit=u, file='cfen'//tag//'.dat', status='replace')
      call write_header(u, 1, finconf)
      write (u, *) '   T      [<energy>]      std.en/sqrt(iconf)      [<spec.heat>]      std.spec.heat/sqrt(iconf)   '
      do it = 1, NTEMP
         T = temperature_for_index(it)
         write (u, '(1x,25(es24.16,1x))') T, cmean(F_EN, it, finconf), cerr(F_2EN, F_EN, it, finconf), &
            cmean(F_SPH, it, finconf), cerr(F_2SPH, F_SPH, it, finconf)
      end do
      close (u)

      ! ---- cfma: magnetisation, susceptibility, Binder
      open (newunit=u, file='cfma'//tag//'.dat', status='replace')
      call write_header(u, 2, finconf)
      write (u, *) '   T      [<mag>]      [<mag>^2]      std.mag/sqrt(iconf)      [<susc>]', &
         ' .     std.susc/sqrt(iconf)      [<Binder>]      std.bin/sqrt(iconf) ', &
         '[log <mag>]  Global.Binder'
      do it = 1, NTEMP
         T = temperature_for_index(it)
         write (u, '(1x,25(e12.6,1x))') T, cmean(F_MAG, it, finconf), cmean(F_2MAG, it, finconf), &
            cerr(F_2MAG, F_MAG, it, finconf), cmean(F_SUSC, it, finconf), cerr(F_2SUSC, F_SUSC, it, finconf), &
            cmean(F_UL, it, finconf), cerr(F_2UL, F_UL, it, finconf), cmean(F_LOGMAG, it, finconf), &
            1 - cmean(F_MAG4, it, finconf)/(3*cmean(F_MAG2, it, finconf)**2)
      end do
      close (u)

      ! ---- cfnm: spatial-mean local nematic
      open (newunit=u, file='cfnm'//tag//'.dat', status='replace')
      call write_header(u, 2, finconf)
      write (u, *) '   T      [<n>]      [|<n>|]      [<|n|>]      [<n>].Error      [|<n>|].Error      '// &
         '[<|n|>].Error      [<n^2>]      [<n^4>]      Global.Nematic.Binder      Nematic.Binder      '// &
         'Nematic.Binder.Error'
      do it = 1, NTEMP
         T = temperature_for_index(it)
         m2 = cmean(F_NBAR2, it, finconf)
         m4 = cmean(F_NBAR4, it, finconf)
         if (m2 > 0.D0) then
            ubn = 1.D0 - m4/(3.D0*m2**2)
         else
            ubn = 0.D0
         end if
         m = cmean(F_NBAR, it, finconf)
         e1 = sqrt(max(0.D0, cmean(F_2NBAR, it, finconf) - m**2))/sqrt(1.D0*finconf)
         v = cfgfields(it, F_CNBAR, 1:finconf)
         mabs = sum(abs(v))/finconf
         e2 = sqrt(max(0.D0, sum((abs(v) - mabs)**2)/finconf))/sqrt(1.D0*finconf)
         v = cfgfields(it, F_CNBARABS, 1:finconf)
         a = cmean(F_NBARABS, it, finconf)
         e3 = sqrt(max(0.D0, sum((v - a)**2)/finconf))/sqrt(1.D0*finconf)
         write (u, '(1x,25(e12.6,1x))') T, m, mabs, a, e1, e2, e3, m2, m4, ubn, cmean(F_UBN, it, finconf), &
            cerr(F_2UBN, F_UBN, it, finconf)
      end do
      close (u)

      ! ---- cfst: stripe order
      open (newunit=u, file='cfst'//tag//'.dat', status='replace')
      call write_header(u, 2, finconf)
      write (u, *) '   T      absstr    str2      str4      nem      Global.Strip.Binder      Strip.susc ', &
         ' Local.Strip.Binder     Local.Strip.Binder.Error      Strip.OP.error      stri.susc.error', &
         ' strx   strx2   strx4   Local.Strx.Binder    Local.Strx.Binder.Error ', &
         ' stry   stry2   stry4   Local.Stry.Binder    Local.Stry.Binder.Error', &
         ' absstrx_absstry  absstrx_absstry.Error  coex_angle  coex_angle.Error'
      do it = 1, NTEMP
         T = temperature_for_index(it)
         s = 0.D0; s2 = 0.D0; q = 0.D0; q2 = 0.D0
         do id = 1, finconf
            a = cfgfields(it, F_STRX, id)*cfgfields(it, F_STRY, id)
            s = s + a
            q = q + coex_angle(cfgfields(it, F_STRX, id), cfgfields(it, F_STRY, id))
            s2 = s2 + a**2
            q2 = q2 + coex_angle(cfgfields(it, F_STRX, id), cfgfields(it, F_STRY, id))**2
         end do
         write (u, '(1x,25(e12.6,1x))') T, cmean(F_ABSSTR, it, finconf), cmean(F_STR2, it, finconf), &
            cmean(F_STR4, it, finconf), cmean(F_NEM, it, finconf), &
            (2 - cmean(F_STR4, it, finconf)/cmean(F_STR2, it, finconf)**2), &
            cmean(F_SUSCSTR, it, finconf), cmean(F_US, it, finconf), cerr(F_2US, F_US, it, finconf), &
            cerr(F_2STR, F_ABSSTR, it, finconf), cerr(F_2SUSCSTR, F_SUSCSTR, it, finconf), &
            cmean(F_STRX, it, finconf), cmean(F_STRX2, it, finconf), cmean(F_STRX4, it, finconf), &
            cmean(F_USX, it, finconf), cerr(F_2USX, F_USX, it, finconf), &
            cmean(F_STRY, it, finconf), cmean(F_STRY2, it, finconf), cmean(F_STRY4, it, finconf), &
            cmean(F_USY, it, finconf), cerr(F_2USY, F_USY, it, finconf), &
            s/finconf, (sqrt(s2/finconf - (s/finconf)**2))/sqrt(1.D0*finconf), &
            q/finconf, (sqrt(q2/finconf - (q/finconf)**2))/sqrt(1.D0*finconf)
      end do
      close (u)

      ! ---- cfhs: energy histogram (bin origin and width of the last config)
      if (WRITEHIST) then
         open (newunit=u, file='cfhs'//tag//'.dat', status='replace')
         call write_header(u, 3, finconf)
         write (u, *) 'T  BIN   [<energy>]  freq'
         do it = 1, NTEMP
            T = temperature_for_index(it)
            do ibin = 1, BINS - 1
               write (u, '(1x, 1(e12.6, 1x), I4, 1x, 2(e12.6, 1x))') T, ibin, &
                  hist_min(it) + hist_bin(it)*(ibin - 1), globalhist(it, ibin)/(finconf)
            end do
         end do
         close (u)
      end if

      ! ---- cfxy: per-config stripe diagnostics
      if (WRITESTRIP) then
         open (newunit=u, file='cfxy'//tag//'.dat', status='replace')
         write (u, *) 'iconf, T, strx, stry, absstr, D, best_en, final_en, absstrx_absstry, coex_angle'
         do id = 1, finconf
            do it = 1, NTEMP
               T = temperature_for_index(it)
               write (u, '(I4,1x,9(E12.6,1x))') id, T, cfgfields(it, F_STRX, id), cfgfields(it, F_STRY, id), &
                  cfgfields(it, F_ABSSTRD, id), cfgfields(it, F_STRDIFF, id), cfgfields(it, F_BESTEN, id), &
                  cfgfields(it, F_FINALEN, id), cfgfields(it, F_STRX, id)*cfgfields(it, F_STRY, id), &
                  coex_angle(cfgfields(it, F_STRX, id), cfgfields(it, F_STRY, id))
            end do
         end do
         close (u)
      end if

      ! ---- cfnd: per-config nematic moments
      open (newunit=u, file='cfnd'//tag//'.dat', status='replace')
      write (u, *) 'iconf, T, nbar, abs_nbar, nbar2, nbar4, nbinder'
      do id = 1, finconf
         do it = 1, NTEMP
            T = temperature_for_index(it)
            write (u, '(I4,1x,6(E12.6,1x))') id, T, cfgfields(it, F_CNBAR, id), cfgfields(it, F_CNBARABS, id), &
               cfgfields(it, F_CNBAR2, id), cfgfields(it, F_CNBAR4, id), cfgfields(it, F_CNBIND, id)
         end do
      end do
      close (u)

      ! ---- cfnu: flips per site
      if (WRITENUMFLIPS) then
         open (newunit=u, file='cfnu'//tag//'.dat', status='replace')
         write (u, *) 'iconf, T, isite, numflips'
         do id = 1, finconf
            do it = 1, NTEMP
               T = temperature_for_index(it)
               do isite = 0, L3 - 1
                  write (u, '(I4,1x,E12.6,1x,I8,1x,I8)') id, T, isite, totalnumflips(isite, it, id)
               end do
            end do
         end do
         close (u)
      end if

      ! ---- cfpt: swap statistics per adjacent pair
      if (SAMPLERMODE == 'PT') then
         open (newunit=u, file='cfpt'//tag//'.dat', status='replace')
         write (u, *) 'pair, T_low, T_high, attempts, accepts, acceptance'
         do it = 1, NTEMP - 1
            if (sum_att(it) > 0) then
               write (u, '(I4,1x,2(E12.6,1x),2(I12,1x),E12.6)') it, temp_ladder(it), temp_ladder(it + 1), &
                  int(sum_att(it), i4b), int(sum_acc(it), i4b), real(sum_acc(it), r8b)/real(sum_att(it), r8b)
            else
               write (u, '(I4,1x,2(E12.6,1x),2(I12,1x),E12.6)') it, temp_ladder(it), temp_ladder(it + 1), &
                  0, 0, 0.D0
            end if
         end do
         write (u, *) 'replicas_visiting_both_extremes ', int(sum_visitboth, i4b), &
            real(sum_visitboth, r8b)/real(max(1, finconf*NTEMP), r8b)
         close (u)
      end if

      ! ---- cfph: the field of every config
      if (WRITEPHITABLE) then
         open (newunit=u, file='cfph'//tag//'.dat', status='replace')
         write (u, *) 'iconf, isite, p'
         do id = 1, finconf
            do isite = 0, L3 - 1
               write (u, '(I4,1x,I8,1x,E12.6)') id, isite, jointp(isite, id)
            end do
         end do
         close (u)
      end if
   end subroutine

end program j1j2_host
