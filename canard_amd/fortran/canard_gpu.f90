!*****
!***** CANARD DRIVER ON THE MI355X STEP (libcnd)
!*****
! Reads the reference's input.canard namelists, builds the multi-block grid
! and initial state, and runs the time loop of src/mo_canard_driver.f90:
! 226-438 with every block's step executed by libcnd on the GPU.  One
! process drives all blocks (one rank per block in the reference).
!
! Namelists: nml_driver, nml_canard, nml_domdcomp, nml_numerics,
! nml_physics, nml_gridgen, nml_sponge as in the reference (README.md:44-99),
! plus nml_gpu for the switches of the GPU library:
!   viscous   (1)  Navier-Stokes (the reference's -DVISCOUS build) or Euler (0)
!   precision (0)  0 FP64, 1 FP32 sweeps with FP64 refinement
!   maxsteps  (-1) stop after this many steps as well (-1: no cap)
!   icase  ('vortex') initial condition for nts=0: 'vortex' (initialo,
!          mo_physics.f90:117-140) or 'tgv' (Taylor-Green vortex)
!   tgvmach (0.1)  Taylor-Green Mach number
!   sponge    (0)  apply the sponge (mo_sponge.f90)
!   graph     (0)  replay whole steps as a hipGraph
! Grids: gridtype 0 (gridtoy) and 1 (gridtoy_multi) of mo_gridgen.f90,
! generated on the device, or gridtype 3: read misc/gridzNNN.dat written by
! the caller (the record layout of mo_io.f90:151-180, one record of (x,y,z)
! per point) and uploaded.
! Output (ndata >= 0): data/grid.plt at set-up and data/nNNN.plt at the
! output times k tmax / ndata, Tecplot files of the recorded averages,
! written by a host thread while the following steps run.

program canard_gpu
   use iso_c_binding
   use mo_cnd
   implicit none

   integer, parameter :: nr = c_double
   real(nr), parameter :: pi = 3.141592653589793_nr, gam = 1.4_nr

   ! nml_driver / nml_canard
   integer :: nio, mbk, ndata, nts, nrestart
   real(nr) :: cfl, tmax
   logical :: ltimer
   ! nml_domdcomp (up to 64 blocks)
   integer, parameter :: mxb = 63
   integer :: nbpc(0:mxb,3), lximb(0:mxb), letmb(0:mxb), lzemb(0:mxb)
   integer :: nbbc(0:mxb,3,2), mbcd(0:mxb,3,2)
   ! nml_numerics / nml_physics / nml_sponge
   real(nr) :: fltk, fltrbc, reoo, tempoo, amach1, amach2, amach3, timf, szco
   integer :: nsmf
   ! nml_gridgen
   integer :: gridtype, lxi0, let0, lze0, lxi1, lxi2, nthick
   real(nr) :: doml0, doml1, domh, span, szth0, szth1, smg, smgvr
   ! nml_gpu
   integer :: viscous, precision, maxsteps, sponge, graph
   character(len=16) :: icase
   real(nr) :: tgvmach

   namelist /nml_driver/ nio, mbk, ndata
   namelist /nml_canard/ nts, nrestart, cfl, tmax, ltimer
   namelist /nml_domdcomp/ nbpc, lximb, letmb, lzemb, nbbc, mbcd
   namelist /nml_numerics/ fltk, fltrbc
   namelist /nml_physics/ reoo, tempoo, amach1, amach2, amach3, timf, nsmf
   namelist /nml_gridgen/ gridtype, lxi0, let0, lze0, doml0, doml1, domh, span, szth0, szth1, nthick, &
                          smg, smgvr, lxi1, lxi2
   namelist /nml_sponge/ szco
   namelist /nml_gpu/ viscous, precision, maxsteps, icase, tgvmach, sponge, graph

   type(cnd_params) :: prm(1)
   type(cnd_block), allocatable :: blk(:)
   type(c_ptr) :: ctx
   real(nr), allocatable :: xyz(:), q(:)
   integer :: mb, nb, npt, n, ndt, nsec, rc
   integer(c_int) :: due, ndati
   real(nr) :: dt, dts, dte, timo, ms, tout
   character(len=32) :: fname
   integer(c_long_long) :: launches
   character(len=8), parameter :: secname(0:7) = [character(len=8) :: 'Filters ', 'TimeStep', 'VSgrad  ', &
                                                   'VSstress', 'Fluxes  ', 'GCBC    ', 'Update  ', 'Average ']

   ! ---- defaults, then the namelists ---------------------------------------------
   nio = 0; mbk = 0; ndata = -1
   nts = 0; nrestart = 0; cfl = 0.8_nr; tmax = 1.0_nr; ltimer = .false.
   nbpc = 1; lximb = 0; letmb = 0; lzemb = 0; nbbc = 45; mbcd = -1
   fltk = 0.49_nr; fltrbc = 0.9_nr
   reoo = 1600.0_nr; tempoo = 300.0_nr; amach1 = 0; amach2 = 0; amach3 = 0; timf = 1; nsmf = 1
   szco = 0; gridtype = 0; lxi0 = 0; let0 = 0; lze0 = 0; lxi1 = 0; lxi2 = 0; nthick = 0
   doml0 = 0; doml1 = 2*pi; domh = pi; span = 2*pi; szth0 = 1; szth1 = 1; smg = 0; smgvr = 0
   viscous = 1; precision = 0; maxsteps = -1; icase = 'vortex'; tgvmach = 0.1_nr; sponge = 0; graph = 0

   call read_nml('nml_driver')
   call read_nml('nml_canard')
   call read_nml('nml_domdcomp')
   call read_nml('nml_numerics')
   call read_nml('nml_physics')
   call read_nml('nml_gridgen')
   call read_nml('nml_sponge')
   call read_nml('nml_gpu')
   if (mbk < 0 .or. mbk > mxb) stop 'Canard: mbk out of range'
   nb = mbk + 1
   do mb = 0, mbk
      if (any(nbpc(mb,:) /= 1)) then
         write(*,*) 'Canard: block', mb, ': nbpc must be 1,1,1 (all blocks of a case live on one GPU)'
         stop 1
      end if
   end do

   ! ---- parameters and blocks -------------------------------------------------------
   prm(1)%fltk = fltk; prm(1)%fltrbc = fltrbc; prm(1)%reoo = reoo; prm(1)%tempoo = tempoo
   prm(1)%amach1 = amach1; prm(1)%amach2 = amach2; prm(1)%amach3 = amach3; prm(1)%timf = timf
   prm(1)%nsmf = nsmf; prm(1)%cfl = cfl; prm(1)%viscous = viscous; prm(1)%sponge = sponge
   prm(1)%szco = szco; prm(1)%szth0 = szth0; prm(1)%szth1 = szth1
   prm(1)%doml0 = doml0; prm(1)%doml1 = doml1
   allocate(blk(nb))
   do mb = 0, mbk
      blk(mb+1)%lxi = lximb(mb); blk(mb+1)%let = letmb(mb); blk(mb+1)%lze = lzemb(mb); blk(mb+1)%inst = 0
      blk(mb+1)%nbc(1,:) = nbbc(mb,:,1); blk(mb+1)%nbc(2,:) = nbbc(mb,:,2)
      blk(mb+1)%mcd(1,:) = mbcd(mb,:,1); blk(mb+1)%mcd(2,:) = mbcd(mb,:,2)
   end do
   write(*,*) 'Canard: initialize the GPU step for', nb, 'block(s)'
   call cnd_check(cnd_create(prm, 1_c_int, blk, int(nb, c_int), int(precision, c_int), ctx), c_null_ptr, 'cnd_create')
   call cnd_check(cnd_use_graph(ctx, int(graph, c_int)), ctx, 'cnd_use_graph')

   ! ---- grid and metrics -------------------------------------------------------------
   write(*,*) 'Canard: grid generation'
   do mb = 0, mbk
      if (gridtype == 0) then          ! gridtoy, on the device
         call cnd_check(cnd_make_grid(ctx, int(mb, c_int), 0_c_int, &
                        [real(lxi0, nr), real(let0, nr), real(lze0, nr), doml0, doml1, domh, span], 7_c_int), &
                        ctx, 'cnd_make_grid')
      else if (gridtype == 1) then     ! gridtoy_multi, on the device
         call cnd_check(cnd_make_grid(ctx, int(mb, c_int), 1_c_int, &
                        [real(lxi0, nr), real(lxi1, nr), real(lxi2, nr), real(let0, nr), real(lze0, nr), &
                         doml0, doml1, domh, span, real(mb, nr)], 10_c_int), ctx, 'cnd_make_grid')
      else                             ! a grid file, uploaded
         npt = (lximb(mb)+1) * (letmb(mb)+1) * (lzemb(mb)+1)
         allocate(xyz(3*npt))
         call make_grid(mb, xyz, npt)
         call cnd_check(cnd_set_grid(ctx, int(mb, c_int), xyz), ctx, 'cnd_set_grid')
         deallocate(xyz)
      end if
   end do
   call cnd_check(cnd_init_metrics(ctx), ctx, 'cnd_init_metrics')

   ! ---- initial conditions (mo_canard_driver.f90:204-214) -------------------------------
   n = 0; ndt = 0; dt = 0; dts = 0; dte = 0; timo = 0
   do mb = 0, mbk
      npt = (lximb(mb)+1) * (letmb(mb)+1) * (lzemb(mb)+1)
      allocate(q(5*npt))
      if (nts == 0) then
         allocate(xyz(3*npt))
         call make_grid(mb, xyz, npt)
         call initial_state(xyz, q, npt)
         deallocate(xyz)
      else
         call read_restart(mb, q, npt, n, ndt, dt, dts, dte, timo)
      end if
      call cnd_check(cnd_set_state(ctx, int(mb, c_int), q), ctx, 'cnd_set_state')
      deallocate(q)
   end do
   call cnd_check(cnd_set_clock(ctx, 0_c_int, int(n, c_int), int(ndt, c_int), dt, dts, dte, timo), ctx, 'cnd_set_clock')

   ! ---- output: grid file, recording and the output times (mo_canard_driver.f90:173-195) --
   if (ndata >= 0) then
      call execute_command_line('mkdir -p data')
      call cnd_check(cnd_output_grid(ctx, 0_c_int, 'data/grid.plt'//c_null_char), ctx, 'cnd_output_grid')
      call cnd_check(cnd_record(ctx, 1_c_int), ctx, 'cnd_record')
      call cnd_check(cnd_output_schedule(ctx, 0_c_int, tmax, int(ndata, c_int)), ctx, 'cnd_output_schedule')
   end if
   if (ltimer) call cnd_check(cnd_profile(ctx, 1_c_int), ctx, 'cnd_profile')

   ! ---- time loop (mo_canard_driver.f90:226-431) ----------------------------------------
   do while (timo < tmax .and. (dt /= 0 .or. n <= 2))
      if (maxsteps >= 0 .and. n >= maxsteps) exit
      write(*,"(' n =',i8,'   time =',f12.5)") n, timo
      call cnd_check(cnd_step(ctx, 1_c_int), ctx, 'cnd_step')
      call cnd_check(cnd_get_clock(ctx, 0_c_int, n, ndt, dt, dts, dte, timo), ctx, 'cnd_get_clock')
      ! output files data/nNNN.plt, written by a host thread while the next
      ! steps run (mo_canard_driver.f90:373-425)
      if (ndata >= 0) then
         call cnd_check(cnd_output_due(ctx, 0_c_int, due, ndati, tout), ctx, 'cnd_output_due')
         if (due == 1) then
            write(fname, "('data/n',i3.3,'.plt')") ndati
            write(*,"(' output ',a,' at time ',f12.5)") trim(fname), tout
            call cnd_check(cnd_output_write(ctx, 0_c_int, trim(fname)//c_null_char, 1_c_int), ctx, 'cnd_output_write')
         end if
      end if
   end do
   if (ndata >= 0) call cnd_check(cnd_output_wait(ctx), ctx, 'cnd_output_wait')

   ! ---- restart files (mo_canard_driver.f90:436-438) ---------------------------------------
   if (nrestart == 1) then
      do mb = 0, mbk
         npt = (lximb(mb)+1) * (letmb(mb)+1) * (lzemb(mb)+1)
         allocate(q(5*npt))
         call cnd_check(cnd_get_state(ctx, int(mb, c_int), q), ctx, 'cnd_get_state')
         call write_restart(mb, q, npt, n, ndt, dt, dts, dte, timo)
         deallocate(q)
      end do
   end if
   if (dt == 0) write(*,*) 'Canard: Overflow.'
   if (ltimer) then
      write(*,'(a)') ' Canard timers (GPU kernel time, ms):'
      do nsec = 0, CND_NSEC - 1
         rc = cnd_profile_read(ctx, int(nsec, c_int), ms, launches)
         write(*,'(3x,a,f14.3,i10)') secname(nsec), ms, launches
      end do
   end if
   call cnd_destroy(ctx)
   write(*,*) 'Canard: Finished.'

contains

   subroutine read_nml(group)
      character(len=*), intent(in) :: group
      integer :: fu, ios
      character(len=256) :: msg
      open(newunit=fu, file='input.canard', action='read', status='old', iostat=ios)
      if (ios /= 0) then
         write(*,*) 'Canard: cannot open input.canard'
         stop 1
      end if
      select case (group)
      case ('nml_driver');   read(fu, nml=nml_driver, iostat=ios, iomsg=msg)
      case ('nml_canard');   read(fu, nml=nml_canard, iostat=ios, iomsg=msg)
      case ('nml_domdcomp'); read(fu, nml=nml_domdcomp, iostat=ios, iomsg=msg)
      case ('nml_numerics'); read(fu, nml=nml_numerics, iostat=ios, iomsg=msg)
      case ('nml_physics');  read(fu, nml=nml_physics, iostat=ios, iomsg=msg)
      case ('nml_gridgen');  read(fu, nml=nml_gridgen, iostat=ios, iomsg=msg)
      case ('nml_sponge');   read(fu, nml=nml_sponge, iostat=ios, iomsg=msg)
      case ('nml_gpu');      read(fu, nml=nml_gpu, iostat=ios, iomsg=msg)
      end select
      close(fu)
      ! a missing group keeps its defaults (end of file); anything else is an error
      if (ios > 0) then
         write(*,*) 'Canard: error reading ', group, ': ', trim(msg)
         stop 1
      end if
   end subroutine read_nml

   ! Grid of block mb (mo_gridgen.f90:510-654, or a grid file).
   subroutine make_grid(mb, xyz, npt)
      integer, intent(in) :: mb, npt
      real(nr), intent(out) :: xyz(npt,3)
      integer :: i, j, k, l, is, js, fu, lxi, let, lze
      real(nr) :: ra0, ra1, ra2, ra3, zs
      lxi = lximb(mb); let = letmb(mb); lze = lzemb(mb)
      select case (gridtype)
      case (0, 1)
         if (gridtype == 0) then
            ra0 = (doml1 + doml0) / lxi0; ra1 = 2.0_nr * domh / let0
            is = 0; js = 0; ra2 = 0; ra3 = 0
         else
            ra0 = (doml1 + doml0) / (lxi0 + lxi1 + lxi2); ra1 = domh / (2.0_nr * let0) * 2.0_nr
            select case (mod(mb, 3))
            case (0); is = 0;               ra2 = 0
            case (1); is = lxi0 + 1;        ra2 = 1
            case (2); is = lxi0 + lxi1 + 2; ra2 = 2
            end select
            if (mb < 3) then
               js = 0; ra3 = 0
            else
               js = let0 + 1; ra3 = 1
            end if
         end if
         l = 0
         do k = 0, lze
            zs = span * (real(lze0 - k, nr) / lze0 - 0.5_nr)
            do j = 0, let
               do i = 0, lxi
                  l = l + 1
                  xyz(l,1) = -doml0 + (i + is - ra2) * ra0
                  xyz(l,2) = -domh + (j + js - ra3) * ra1
                  xyz(l,3) = zs
               end do
            end do
         end do
      case (3)
         open(newunit=fu, file=grid_name(mb), access='direct', form='unformatted', recl=24, status='old')
         do l = 1, npt
            read(fu, rec=l) xyz(l,1), xyz(l,2), xyz(l,3)
         end do
         close(fu)
      case default
         write(*,*) 'Canard: unsupported gridtype', gridtype
         stop 1
      end select
   end subroutine make_grid

   function grid_name(mb) result(s)
      integer, intent(in) :: mb
      character(len=32) :: s
      write(s, '(a,i3.3,a)') 'misc/gridz', mb, '.dat'
   end function grid_name

   function restart_name(mb) result(s)
      integer, intent(in) :: mb
      character(len=32) :: s
      write(s, '(a,i3.3,a)') 'misc/restartz', mb, '.dat'
   end function restart_name

   ! nts = 0 initial condition
   subroutine initial_state(xyz, q, npt)
      integer, intent(in) :: npt
      real(nr), intent(in) :: xyz(npt,3)
      real(nr), intent(out) :: q(npt,5)
      integer :: l
      real(nr) :: x, y, z, u, v, p, rho, ao, bo, ve(3), hv2
      real(nr), parameter :: radv = 1, k1 = 12.5_nr, k2 = 1
      do l = 1, npt
         x = xyz(l,1); y = xyz(l,2); z = xyz(l,3)
         if (trim(icase) == 'tgv') then
            u = tgvmach * sin(x) * cos(y) * cos(z)
            v = -tgvmach * cos(x) * sin(y) * cos(z)
            p = 1 / gam + tgvmach**2 / 16 * (cos(2*x) + cos(2*y)) * (cos(2*z) + 2)
            rho = gam * p
            q(l,1) = rho; q(l,2) = rho*u; q(l,3) = rho*v; q(l,4) = 0
            q(l,5) = p / (gam - 1) + 0.5_nr * rho * (u*u + v*v)
         else
            ! isentropic vortex of initialo (mo_physics.f90:126-138)
            ao = k2 / 2.0_nr / pi * sqrt(exp(1 - k1**2 * (x**2 + y**2) / radv**2))
            bo = (1 - 0.5_nr * (gam - 1) * ao * ao)**(1 / (gam - 1))
            ve = [k1 * y * ao / radv, -k1 * x * ao / radv, 0.0_nr]
            hv2 = 0.5_nr * (ve(1)**2 + ve(2)**2 + ve(3)**2)
            q(l,1) = bo
            q(l,2:4) = bo * ve
            q(l,5) = (1 / gam) / (gam - 1) * bo**gam + hv2 * bo
         end if
      end do
   end subroutine initial_state

   ! Restart file of block mb (mo_io.f90:184-271): record 1 (n, ndt, dt, dts,
   ! dte), record 2 (timo, 0, 0, 0, 0), then q of each point, i fastest.
   subroutine read_restart(mb, q, npt, n, ndt, dt, dts, dte, timo)
      integer, intent(in) :: mb, npt
      real(nr), intent(out) :: q(npt,5)
      integer, intent(out) :: n, ndt
      real(nr), intent(out) :: dt, dts, dte, timo
      real(nr) :: cha(5), dha(5)
      integer :: fu, l, ios
      open(newunit=fu, file=restart_name(mb), access='direct', form='unformatted', recl=40, status='old', iostat=ios)
      if (ios /= 0) then
         write(*,*) 'Canard: cannot open ', trim(restart_name(mb))
         stop 1
      end if
      read(fu, rec=1) cha
      read(fu, rec=2) dha
      n = nint(cha(1)); ndt = nint(cha(2)); dt = cha(3); dts = cha(4); dte = cha(5); timo = dha(1)
      do l = 1, npt
         read(fu, rec=l+2, iostat=ios) q(l,:)
         if (ios /= 0) then
            write(*,*) 'Canard: truncated restart file ', trim(restart_name(mb))
            stop 1
         end if
      end do
      close(fu)
   end subroutine read_restart

   subroutine write_restart(mb, q, npt, n, ndt, dt, dts, dte, timo)
      integer, intent(in) :: mb, npt, n, ndt
      real(nr), intent(in) :: q(npt,5)
      real(nr), intent(in) :: dt, dts, dte, timo
      integer :: fu, l
      open(newunit=fu, file=restart_name(mb), access='direct', form='unformatted', recl=40, status='replace')
      write(fu, rec=1) real(n, nr), real(ndt, nr), dt, dts, dte
      write(fu, rec=2) timo, 0.0_nr, 0.0_nr, 0.0_nr, 0.0_nr
      do l = 1, npt
         write(fu, rec=l+2) q(l,:)
      end do
      close(fu)
   end subroutine write_restart

end program canard_gpu
