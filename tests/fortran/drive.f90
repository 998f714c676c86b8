!! Host-side driver of the Fortran module `vilma` for the tests:
!!   drive <namelist> --print-params [defaults]  (CPU only: vilma_par_load result)
!!   drive <namelist> <input.bin> <output.bin> [--host-grid] [--spinup]  (GPU: init + update loop)
!! input.bin (stream): int32 nphi, nlat, nsteps; real64 dt_yr; z_bed_eq(nphi,nlat);
!! h_ice(nphi,nlat,nsteps). output.bin: rsl, z_bed, C (nphi,nlat), bsl, time.
!! With --host-grid the fields are on a lon-lat host grid: nphi, nlat are its
!! dimensions, and its axes lon(nphi), lat(nlat) [degrees] follow dt_yr; rsl
!! and z_bed come back on it (C stays on the Gauss grid). With --spinup,
!! solid_earth_spinup under the first ice field precedes the loop (the
!! namelist's equil_time_max, equil_rate_tol, pre_spinup_1d).
program drive
   use vilma
   implicit none
   type(solid_earth) :: se
   type(grid_class) :: grid
   character(len=512) :: nml, a2, a3, a4, a5
   integer :: nphi, nlat, nsteps, k, u, i
   real(wp) :: dt
   real(wp), allocatable :: z(:,:), hice(:,:,:), heq(:,:), lon(:), lat(:)
   logical :: host, spin

   call get_command_argument(1, nml)
   call get_command_argument(2, a2)
   call get_command_argument(3, a3)
   call get_command_argument(4, a4)
   call get_command_argument(5, a5)
   host = trim(a4) == "--host-grid" .or. trim(a5) == "--host-grid"
   spin = trim(a4) == "--spinup" .or. trim(a5) == "--spinup"
   if (trim(a2) == "--print-params" .and. len_trim(a3) > 0) then
      call vilma_par_load(se%par, trim(nml), defaults_file=trim(a3))
   else
      call vilma_par_load(se%par, trim(nml))
   end if
   if (trim(a2) == "--print-params") then
      write(*, '(a,5(1x,i0))') 'grid', se%par%lmax, se%par%nlat, se%par%nphi, se%par%mmax, se%par%mres
      write(*, '(a,1x,a,1x,i0)') 'earth', trim(se%par%earth), se%par%n_layer
      write(*, '(a,16(1x,es24.16e3))') 'r_bot', (se%par%r_bot(i), i = 1, 16)
      write(*, '(a,16(1x,es24.16e3))') 'eta', (se%par%eta(i), i = 1, 16)
      write(*, '(a,16(1x,i0))') 'rheology', (se%par%rheology(i), i = 1, 16)
      write(*, '(a,3(1x,es24.16e3))') 'dt', se%par%dt_init, se%par%dt_min, se%par%dt_max
      write(*, '(a,1x,es24.16e3,2(1x,i0))') 'sle', se%par%sle_tol, se%par%sle_n_outer, se%par%sle_n_inner
      write(*, '(a,2(1x,l1),1x,a,1x,a)') 'flags', se%par%rotation, se%par%sle_subgrid, trim(se%par%deg1_frame), &
         trim(se%par%scheme)
      write(*, '(a,2(1x,es24.16e3))') 'tol', se%par%visc3d_tol, se%par%skip_tol
      stop
   end if
   open(newunit=u, file=trim(a2), access='stream', form='unformatted', status='old', action='read')
   read(u) nphi, nlat, nsteps
   read(u) dt
   if (host) then
      allocate(lon(nphi), lat(nlat))
      read(u) lon
      read(u) lat
      call grid_init(grid, name="host", mtype="latlon", units="degrees", x=lon, y=lat)
   end if
   allocate(z(nphi, nlat), heq(nphi, nlat), hice(nphi, nlat, nsteps))
   read(u) z
   read(u) hice
   close(u)
   heq = 0.0_wp
   if (host) then
      call solid_earth_init(se, z, heq, grid=grid)
   else
      call solid_earth_init(se, z, heq)
   end if
   if (spin) call solid_earth_spinup(se, hice(:, :, 1))
   do k = 1, nsteps
      call solid_earth_update(se, hice(:, :, k), dt)
   end do
   open(newunit=u, file=trim(a3), access='stream', form='unformatted', status='replace', action='write')
   write(u) se%rsl
   write(u) se%z_bed
   write(u) se%gg%C
   write(u) se%bsl, se%time
   close(u)
   call solid_earth_finalize(se)
end program drive
