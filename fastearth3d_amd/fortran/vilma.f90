!! vilma.f90 -- Fortran host module `vilma` over the MI355X C-ABI (include/vilma_hip.h).
!!
!! A drop-in for the reference's `use vilma` (src/vilma.f90:6-19) on the path a
!! host such as CLIMBER-X drives: the parameter record and its namelist loader
!! (vilma_param_class / vilma_par_load, src/vilma_params.f90:26-205, :209-331),
!! and the solid_earth type with init / update / finalize and the lateral
!! viscosity switch (src/vilma_coupling.f90:71-140, :144, :320, :336, :599).
!! Same names, argument order and output components; a nonzero status from the
!! library becomes `error stop <message>` as in the reference.
!!
!! Differences at this boundary: the optional `grid` argument of
!! solid_earth_init takes this module's lon-lat look-alike of
!! coords::grid_class (grid_init); the lateral viscosity of
!! solid_earth_enable_visc_3d comes in as an array (a generic interface over
!! the per-element perturbation (nphi, nlat, ne) and the nodal absolute
!! log10 viscosity (nphi*nlat, nr)) instead of being read from netCDF;
!! vilma_par_load reads the group "vilma" with the intrinsic NAMELIST facility
!! (overlay of `filename` on `defaults_file`, dt_* and equil_time_max in
!! years converted to seconds).
module vilma
   use, intrinsic :: iso_c_binding
   implicit none
   private

   public :: wp, MAX_LAYER, vilma_param_class, vilma_par_load, vilma_par_print
   public :: solid_earth, solid_earth_init, solid_earth_update, solid_earth_finalize
   public :: solid_earth_enable_visc_3d, solid_earth_enable_visc_3d_field, solid_earth_enable_visc_3d_nodes
   public :: solid_earth_spinup, solid_earth_sync_host
   public :: solid_earth_get_restart_state, solid_earth_set_restart_state, solid_earth_set_restart_memory
   public :: remap_ll_gauss, remap_init, remap_to_gauss, remap_to_ll, remap_free
   public :: grid_class, grid_init

   integer, parameter :: wp = c_double
   integer, parameter :: MAX_LAYER = 16
   real(wp), parameter :: sec_per_year = 31556926.0_wp

   !! the parameter record (field names and defaults of src/vilma_params.f90:26-205)
   type :: vilma_param_class
      character(len=16) :: solver = "v2"
      integer  :: lmax = 0, nlat = 0, nphi = 0, mmax = -1, mres = 1
      real(wp) :: eps_polar = -1.0_wp
      character(len=64) :: earth = "M3-L70-V01"
      integer  :: n_layer = 0
      real(wp) :: r_earth = 6371.0e3_wp, r_core = 3480.0e3_wp
      real(wp) :: r_bot(MAX_LAYER) = 0.0_wp, r_top(MAX_LAYER) = 0.0_wp
      real(wp) :: rho(MAX_LAYER) = 0.0_wp, mu(MAX_LAYER) = 0.0_wp, eta(MAX_LAYER) = 0.0_wp
      integer  :: rheology(MAX_LAYER) = 1
      integer  :: n_elem_layer(MAX_LAYER) = 0       !! additive: explicit elements per layer
      character(len=8) :: scheme = "fe"
      integer  :: max_couple_iter = 20
      character(len=8) :: earth_response = "ve"
      integer  :: sle_n_outer = 3, sle_n_inner = 20
      real(wp) :: sle_tol = 1.0e-7_wp
      integer  :: sle_max_mem_iter = 20
      logical  :: sle_fixed_ocean = .false., sle_subgrid = .true.
      real(wp) :: dt_init = 0.0_wp, dt_min = 0.0_wp, dt_max = huge(1.0_wp)   !! seconds
      real(wp) :: rtol = 1.0e-4_wp, atol = 1.0e-3_wp, safety = 0.9_wp
      real(wp) :: grow_max = 5.0_wp, shrink_min = 0.2_wp, cfl = 1.0_wp
      logical  :: rotation = .true.
      real(wp) :: equil_time_max = 0.0_wp, equil_rate_tol = 1.0e-3_wp
      logical  :: pre_spinup_1d = .false.
      character(len=8) :: deg1_frame = "cm"
      logical  :: l_visc_3d = .false.
      character(len=512) :: visc_3d_file = ""
      character(len=64) :: name_visc = "eta", name_visc_lon = "lon", name_visc_lat = "lat"
      character(len=64) :: name_visc_r = "r", name_visc_sd = ""
      real(wp) :: f_visc_sd = 0.0_wp, f_visc_rel = 0.1_wp
      real(wp) :: visc_log10_min = 19.5_wp, visc_log10_max = 30.0_wp
      real(wp) :: visc3d_lid_depth = 0.0_wp, visc3d_lid_log10max = 22.0_wp
      real(wp) :: visc3d_tol = 1.0e-3_wp
      logical  :: l_toroidal = .false.
      real(wp) :: skip_tol = 1.0e-4_wp                 !! additive: drift skip gate
      ! VILMA-v1 backend settings: accepted so reference namelists load, inert here
      integer  :: vilma_v1_jmax = 170, vilma_v1_l_prem = 1, vilma_v1_nsub = 1
      character(len=512) :: vilma_v1_input_dir = "input/vilma", vilma_v1_out_dir = "vilma_v1"
      character(len=512) :: vilma_v1_grid_file = "input/vilma/vilma_grid.nc"
      character(len=128) :: vilma_v1_visc_1d_file = "visko.inp", vilma_v1_visc_3d_file = "visc3d_Bagge2021.nc"
   end type vilma_param_class

   !! C mirror of vh_params (include/vilma_hip.h)
   type, bind(c) :: vh_params_c
      integer(c_int) :: lmax, nlat, nphi, mmax, mres
      real(c_double) :: eps_polar
      character(kind=c_char) :: earth(64)
      integer(c_int) :: n_layer
      real(c_double) :: r_earth, r_core
      real(c_double) :: r_bot(MAX_LAYER), r_top(MAX_LAYER)
      real(c_double) :: rho(MAX_LAYER), mu(MAX_LAYER), eta(MAX_LAYER)
      integer(c_int) :: rheology(MAX_LAYER)
      integer(c_int) :: n_elem_layer(MAX_LAYER)
      character(kind=c_char) :: scheme(8)
      integer(c_int) :: max_couple_iter
      character(kind=c_char) :: earth_response(8)
      integer(c_int) :: sle_n_outer, sle_n_inner
      real(c_double) :: sle_tol
      integer(c_int) :: sle_max_mem_iter, sle_fixed_ocean, sle_subgrid
      real(c_double) :: dt_init, dt_min, dt_max
      real(c_double) :: rtol, atol, safety, grow_max, shrink_min, cfl
      integer(c_int) :: rotation
      character(kind=c_char) :: deg1_frame(8)
      integer(c_int) :: l_visc_3d
      real(c_double) :: visc3d_tol
      integer(c_int) :: l_toroidal
      real(c_double) :: visc_log10_min, visc_log10_max, visc3d_lid_depth, visc3d_lid_log10max
      real(c_double) :: skip_tol
   end type vh_params_c

   !! the Gauss-grid state the host reads (se%gg%..., src/vilma_coupling.f90:71-140)
   type :: gauss_grid_fields
      real(wp), allocatable :: rsl(:,:), z_bed(:,:), C(:,:), C0(:,:), h_ice(:,:), s_rot(:,:)
   end type gauss_grid_fields

   !! A map pair between a host lon-lat grid and a model's Gauss grid
   !! (remap_ll_gauss, src/vilma_remap.f90:44-54): conservative in, bilinear out.
   type :: remap_ll_gauss
      type(c_ptr) :: m = c_null_ptr
      integer :: nlon = 0, nlat_ll = 0        !! host dimensions
      integer :: nphi = 0, nlat = 0           !! Gauss dimensions
   end type remap_ll_gauss

   !! The part of fesm-utils coords::grid_class the coupling reads (grid%G%x,
   !! grid%G%y: lon-lat cell-centre axes in degrees, ascending), filled by
   !! grid_init; a host passes it to solid_earth_init as it would the coords one.
   type :: grid_axes
      integer :: nx = 0, ny = 0
      real(wp), allocatable :: x(:), y(:)
   end type grid_axes
   type :: grid_class
      character(len=256) :: name = ""
      type(grid_axes) :: G
   end type grid_class

   type :: solid_earth
      type(vilma_param_class) :: par
      real(wp), allocatable :: rsl(:,:), z_bed(:,:)
      real(wp) :: bsl = 0.0_wp, time = 0.0_wp, worst_mass_resid = 0.0_wp
      complex(wp), allocatable :: sigma_lm(:)
      type(gauss_grid_fields) :: gg
      type(c_ptr) :: h = c_null_ptr
      integer :: nphi = 0, nlat = 0
      ! host-grid coupling (src/vilma_coupling.f90:115-117): the reference fields
      ! at host resolution and the map pair, when the host grid is not the Gauss grid
      logical :: remap = .false.
      type(remap_ll_gauss) :: map
      ! the reference state and the current ice at host resolution (src/vilma_coupling.f90:131-133)
      real(wp), allocatable :: z_bed_eq(:,:), h_ice_eq(:,:), h_ice(:,:)
   end type solid_earth

   !! solid_earth_enable_visc_3d (src/vilma_coupling.f90:320-334) with the
   !! lateral field given in memory, per element or per node
   interface solid_earth_enable_visc_3d
      module procedure solid_earth_enable_visc_3d_field, solid_earth_enable_visc_3d_nodes
   end interface solid_earth_enable_visc_3d


   interface
      integer(c_int) function vh_create(p, nphi, nlat, z, hi, hi0, out) bind(c, name="vh_create")
         import :: c_int, c_ptr, vh_params_c
         type(vh_params_c), intent(in) :: p
         integer(c_int), value :: nphi, nlat
         type(c_ptr), value :: z, hi, hi0
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function vh_update(h, hi, dt_yr) bind(c, name="vh_update")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, hi
         real(c_double), value :: dt_yr
      end function
      integer(c_int) function vh_get_field(h, which, out) bind(c, name="vh_get_field")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, out
         integer(c_int), value :: which
      end function
      integer(c_int) function vh_get_scalar(h, which, out) bind(c, name="vh_get_scalar")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         integer(c_int), value :: which
         real(c_double), intent(out) :: out
      end function
      integer(c_int) function vh_get_spectrum(h, which, out) bind(c, name="vh_get_spectrum")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, out
         integer(c_int), value :: which
      end function
      integer(c_int) function vh_get_dims(h, which, out) bind(c, name="vh_get_dims")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: which
         integer(c_int), intent(out) :: out
      end function
      integer(c_int) function vh_enable_visc_3d(h, pert) bind(c, name="vh_enable_visc_3d")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, pert
      end function
      integer(c_int) function vh_enable_visc_3d_nodes(h, visc, lid_depth, lid_max, cap) &
            bind(c, name="vh_enable_visc_3d_nodes")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, visc
         real(c_double), value :: lid_depth, lid_max, cap
      end function
      integer(c_int) function vh_get_memory(h, tau) bind(c, name="vh_get_memory")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, tau
      end function
      integer(c_int) function vh_set_memory(h, tau) bind(c, name="vh_set_memory")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, tau
      end function
      integer(c_int) function vh_set_memory_lowres(h, tau, nk_src) bind(c, name="vh_set_memory_lowres")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, tau
         integer(c_int), value :: nk_src
      end function
      integer(c_int) function vh_get_step_state(h, dt_try, primed, sig) bind(c, name="vh_get_step_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, sig
         real(c_double) :: dt_try
         integer(c_int) :: primed
      end function
      integer(c_int) function vh_set_step_state(h, dt_try, primed, sig) bind(c, name="vh_set_step_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, sig
         real(c_double), value :: dt_try
         integer(c_int), value :: primed
      end function
      integer(c_int) function vh_get_state(h, t, rsl, hi, m, lmem, tmem) bind(c, name="vh_get_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, rsl, hi, m, lmem, tmem
         real(c_double), intent(out) :: t
      end function
      integer(c_int) function vh_set_state(h, t, rsl, hi, m, lmem, tmem) bind(c, name="vh_set_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, rsl, hi, m, lmem, tmem
         real(c_double), value :: t
      end function
      integer(c_int) function vh_remap_create(nphi, nlat, nlon, nlat_ll, lon, lat, out) &
            bind(c, name="vh_remap_create")
         import :: c_int, c_ptr
         integer(c_int), value :: nphi, nlat, nlon, nlat_ll
         type(c_ptr), value :: lon, lat
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function vh_remap_to_gauss(m, f, g, cm) bind(c, name="vh_remap_to_gauss")
         import :: c_int, c_ptr
         type(c_ptr), value :: m, f, g
         integer(c_int), value :: cm
      end function
      integer(c_int) function vh_remap_to_ll(m, g, f) bind(c, name="vh_remap_to_ll")
         import :: c_int, c_ptr
         type(c_ptr), value :: m, g, f
      end function
      integer(c_int) function vh_remap_destroy(m) bind(c, name="vh_remap_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: m
      end function
      integer(c_int) function vh_spinup(h, ice, tmax, tol, pre1d) bind(c, name="vh_spinup")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h, ice
         real(c_double), value :: tmax, tol
         integer(c_int), value :: pre1d
      end function
      integer(c_int) function vh_destroy(h) bind(c, name="vh_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      type(c_ptr) function vh_last_error(h) bind(c, name="vh_last_error")
         import :: c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   !! vilma_par_load (src/vilma_params.f90:209-331): `filename`'s &vilma group
   !! overlaid on `defaults_file`'s (each defaults to the other); dt_init,
   !! dt_min, dt_max and equil_time_max are years in the file, seconds in p.
   subroutine vilma_par_load(p, filename, defaults_file, group)
      type(vilma_param_class), intent(inout) :: p
      character(len=*), intent(in) :: filename
      character(len=*), intent(in), optional :: defaults_file, group
      character(len=16) :: solver
      integer  :: lmax, nlat, nphi, mmax, mres, n_layer, rheology(MAX_LAYER), n_elem_layer(MAX_LAYER)
      integer  :: max_couple_iter, sle_n_outer, sle_n_inner, sle_max_mem_iter
      integer  :: vilma_v1_jmax, vilma_v1_l_prem, vilma_v1_nsub
      real(wp) :: eps_polar, r_earth, r_core, r_bot(MAX_LAYER), r_top(MAX_LAYER)
      real(wp) :: rho(MAX_LAYER), mu(MAX_LAYER), eta(MAX_LAYER), sle_tol
      real(wp) :: dt_init, dt_min, dt_max, rtol, atol, safety, grow_max, shrink_min, cfl
      real(wp) :: equil_time_max, equil_rate_tol, f_visc_sd, f_visc_rel, visc_log10_min, visc_log10_max
      real(wp) :: visc3d_lid_depth, visc3d_lid_log10max, visc3d_tol, skip_tol
      character(len=64) :: earth, name_visc, name_visc_lon, name_visc_lat, name_visc_r, name_visc_sd
      character(len=8)  :: scheme, earth_response, deg1_frame
      character(len=512) :: visc_3d_file, vilma_v1_input_dir, vilma_v1_out_dir, vilma_v1_grid_file
      character(len=128) :: vilma_v1_visc_1d_file, vilma_v1_visc_3d_file
      logical  :: sle_fixed_ocean, sle_subgrid, rotation, pre_spinup_1d, l_visc_3d, l_toroidal
      namelist /vilma/ solver, lmax, nlat, nphi, mmax, mres, eps_polar, earth, n_layer, r_earth, r_core, &
         r_bot, r_top, rho, mu, eta, rheology, n_elem_layer, scheme, max_couple_iter, earth_response, &
         sle_n_outer, sle_n_inner, sle_tol, sle_max_mem_iter, sle_fixed_ocean, sle_subgrid, &
         dt_init, dt_min, dt_max, rtol, atol, safety, grow_max, shrink_min, cfl, rotation, &
         equil_time_max, equil_rate_tol, pre_spinup_1d, deg1_frame, l_visc_3d, visc_3d_file, &
         name_visc, name_visc_lon, name_visc_lat, name_visc_r, name_visc_sd, f_visc_sd, f_visc_rel, &
         visc_log10_min, visc_log10_max, visc3d_lid_depth, visc3d_lid_log10max, visc3d_tol, &
         l_toroidal, skip_tol, vilma_v1_jmax, vilma_v1_input_dir, vilma_v1_out_dir, vilma_v1_grid_file, &
         vilma_v1_visc_1d_file, vilma_v1_visc_3d_file, vilma_v1_l_prem, vilma_v1_nsub

      if (present(group)) then
         if (trim(group) /= "vilma") error stop 'vilma_par_load: only the namelist group "vilma" is supported'
      end if
      solver = p%solver; lmax = p%lmax; nlat = p%nlat; nphi = p%nphi; mmax = p%mmax; mres = p%mres
      eps_polar = p%eps_polar; earth = p%earth; n_layer = p%n_layer; r_earth = p%r_earth; r_core = p%r_core
      r_bot = p%r_bot; r_top = p%r_top; rho = p%rho; mu = p%mu; eta = p%eta; rheology = p%rheology
      n_elem_layer = p%n_elem_layer; scheme = p%scheme; max_couple_iter = p%max_couple_iter
      earth_response = p%earth_response; sle_n_outer = p%sle_n_outer; sle_n_inner = p%sle_n_inner
      sle_tol = p%sle_tol; sle_max_mem_iter = p%sle_max_mem_iter; sle_fixed_ocean = p%sle_fixed_ocean
      sle_subgrid = p%sle_subgrid
      dt_init = p%dt_init / sec_per_year; dt_min = p%dt_min / sec_per_year; dt_max = p%dt_max / sec_per_year
      rtol = p%rtol; atol = p%atol; safety = p%safety; grow_max = p%grow_max; shrink_min = p%shrink_min
      cfl = p%cfl; rotation = p%rotation
      equil_time_max = p%equil_time_max / sec_per_year; equil_rate_tol = p%equil_rate_tol
      pre_spinup_1d = p%pre_spinup_1d; deg1_frame = p%deg1_frame; l_visc_3d = p%l_visc_3d
      visc_3d_file = p%visc_3d_file; name_visc = p%name_visc; name_visc_lon = p%name_visc_lon
      name_visc_lat = p%name_visc_lat; name_visc_r = p%name_visc_r; name_visc_sd = p%name_visc_sd
      f_visc_sd = p%f_visc_sd; f_visc_rel = p%f_visc_rel; visc_log10_min = p%visc_log10_min
      visc_log10_max = p%visc_log10_max; visc3d_lid_depth = p%visc3d_lid_depth
      visc3d_lid_log10max = p%visc3d_lid_log10max; visc3d_tol = p%visc3d_tol; l_toroidal = p%l_toroidal
      skip_tol = p%skip_tol; vilma_v1_jmax = p%vilma_v1_jmax; vilma_v1_input_dir = p%vilma_v1_input_dir
      vilma_v1_out_dir = p%vilma_v1_out_dir; vilma_v1_grid_file = p%vilma_v1_grid_file
      vilma_v1_visc_1d_file = p%vilma_v1_visc_1d_file; vilma_v1_visc_3d_file = p%vilma_v1_visc_3d_file
      vilma_v1_l_prem = p%vilma_v1_l_prem; vilma_v1_nsub = p%vilma_v1_nsub

      if (present(defaults_file)) call read_group(defaults_file)
      call read_group(filename)

      if (trim(deg1_frame) /= "cf" .and. trim(deg1_frame) /= "cm") &
         error stop 'vilma_params: deg1_frame must be "cf" or "cm"'
      if (vilma_v1_l_prem /= 0 .and. vilma_v1_l_prem /= 1) error stop 'vilma_params: vilma_v1_l_prem must be 0 or 1'
      if (vilma_v1_nsub < 1) error stop 'vilma_params: vilma_v1_nsub must be >= 1'

      p%solver = solver; p%lmax = lmax; p%nlat = nlat; p%nphi = nphi; p%mmax = mmax; p%mres = mres
      p%eps_polar = eps_polar; p%earth = earth; p%n_layer = n_layer; p%r_earth = r_earth; p%r_core = r_core
      p%r_bot = r_bot; p%r_top = r_top; p%rho = rho; p%mu = mu; p%eta = eta; p%rheology = rheology
      p%n_elem_layer = n_elem_layer; p%scheme = scheme; p%max_couple_iter = max_couple_iter
      p%earth_response = earth_response; p%sle_n_outer = sle_n_outer; p%sle_n_inner = sle_n_inner
      p%sle_tol = sle_tol; p%sle_max_mem_iter = sle_max_mem_iter; p%sle_fixed_ocean = sle_fixed_ocean
      p%sle_subgrid = sle_subgrid
      p%dt_init = dt_init * sec_per_year; p%dt_min = dt_min * sec_per_year; p%dt_max = dt_max * sec_per_year
      p%rtol = rtol; p%atol = atol; p%safety = safety; p%grow_max = grow_max; p%shrink_min = shrink_min
      p%cfl = cfl; p%rotation = rotation
      p%equil_time_max = equil_time_max * sec_per_year; p%equil_rate_tol = equil_rate_tol
      p%pre_spinup_1d = pre_spinup_1d; p%deg1_frame = deg1_frame; p%l_visc_3d = l_visc_3d
      p%visc_3d_file = visc_3d_file; p%name_visc = name_visc; p%name_visc_lon = name_visc_lon
      p%name_visc_lat = name_visc_lat; p%name_visc_r = name_visc_r; p%name_visc_sd = name_visc_sd
      p%f_visc_sd = f_visc_sd; p%f_visc_rel = f_visc_rel; p%visc_log10_min = visc_log10_min
      p%visc_log10_max = visc_log10_max; p%visc3d_lid_depth = visc3d_lid_depth
      p%visc3d_lid_log10max = visc3d_lid_log10max; p%visc3d_tol = visc3d_tol; p%l_toroidal = l_toroidal
      p%skip_tol = skip_tol; p%vilma_v1_jmax = vilma_v1_jmax; p%vilma_v1_input_dir = vilma_v1_input_dir
      p%vilma_v1_out_dir = vilma_v1_out_dir; p%vilma_v1_grid_file = vilma_v1_grid_file
      p%vilma_v1_visc_1d_file = vilma_v1_visc_1d_file; p%vilma_v1_visc_3d_file = vilma_v1_visc_3d_file
      p%vilma_v1_l_prem = vilma_v1_l_prem; p%vilma_v1_nsub = vilma_v1_nsub

   contains

      subroutine read_group(fname)
         character(len=*), intent(in) :: fname
         integer :: u, ios
         character(len=256) :: msg
         open(newunit=u, file=fname, status="old", action="read", iostat=ios, iomsg=msg)
         if (ios /= 0) then
            write(*, '(a)') 'vilma_par_load: cannot open ' // trim(fname) // ': ' // trim(msg)
            error stop 'vilma_par_load: cannot open the parameter file'
         end if
         read(u, nml=vilma, iostat=ios, iomsg=msg)
         close(u)
         if (ios /= 0) then
            write(*, '(a)') 'vilma_par_load: ' // trim(fname) // ': ' // trim(msg)
            error stop 'vilma_par_load: bad &vilma group'
         end if
      end subroutine read_group

   end subroutine vilma_par_load

   subroutine put_str(dst, src)
      character(kind=c_char), intent(out) :: dst(:)
      character(len=*), intent(in) :: src
      integer :: i, n
      dst = c_null_char
      n = min(len_trim(src), size(dst) - 1)
      do i = 1, n
         dst(i) = src(i:i)
      end do
   end subroutine put_str

   function to_c(p) result(c)
      type(vilma_param_class), intent(in) :: p
      type(vh_params_c) :: c
      c%lmax = p%lmax; c%nlat = p%nlat; c%nphi = p%nphi; c%mmax = p%mmax; c%mres = p%mres
      c%eps_polar = p%eps_polar
      call put_str(c%earth, p%earth)
      c%n_layer = p%n_layer; c%r_earth = p%r_earth; c%r_core = p%r_core
      c%r_bot = p%r_bot; c%r_top = p%r_top; c%rho = p%rho; c%mu = p%mu; c%eta = p%eta
      c%rheology = p%rheology; c%n_elem_layer = p%n_elem_layer
      call put_str(c%scheme, p%scheme)
      c%max_couple_iter = p%max_couple_iter
      call put_str(c%earth_response, p%earth_response)
      c%sle_n_outer = p%sle_n_outer; c%sle_n_inner = p%sle_n_inner; c%sle_tol = p%sle_tol
      c%sle_max_mem_iter = p%sle_max_mem_iter
      c%sle_fixed_ocean = merge(1, 0, p%sle_fixed_ocean); c%sle_subgrid = merge(1, 0, p%sle_subgrid)
      c%dt_init = p%dt_init; c%dt_min = p%dt_min; c%dt_max = p%dt_max
      c%rtol = p%rtol; c%atol = p%atol; c%safety = p%safety; c%grow_max = p%grow_max
      c%shrink_min = p%shrink_min; c%cfl = p%cfl
      c%rotation = merge(1, 0, p%rotation)
      call put_str(c%deg1_frame, p%deg1_frame)
      c%l_visc_3d = merge(1, 0, p%l_visc_3d); c%visc3d_tol = p%visc3d_tol
      c%l_toroidal = merge(1, 0, p%l_toroidal)
      c%visc_log10_min = p%visc_log10_min; c%visc_log10_max = p%visc_log10_max
      c%visc3d_lid_depth = p%visc3d_lid_depth; c%visc3d_lid_log10max = p%visc3d_lid_log10max
      c%skip_tol = p%skip_tol
   end function to_c

   !! nonzero status -> error stop with the library's message (the reference's convention)
   subroutine check(self, rc, what)
      type(solid_earth), intent(in) :: self
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      type(c_ptr) :: m
      character(kind=c_char), pointer :: s(:)
      character(len=:), allocatable :: msg
      integer :: i, n
      if (rc == 0) return
      m = vh_last_error(self%h)
      n = int(c_strlen(m))
      call c_f_pointer(m, s, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = s(i)
      end do
      write(*, '(a)') trim(what) // ': ' // msg
      error stop 'vilma: ' // what // ' failed (message above)'
   end subroutine check

   subroutine refresh(self)
      type(solid_earth), intent(inout), target :: self
      real(wp) :: v
      call check(self, vh_get_field(self%h, 0, c_loc(self%gg%rsl)), "solid_earth_update")
      call check(self, vh_get_field(self%h, 1, c_loc(self%gg%z_bed)), "solid_earth_update")
      call check(self, vh_get_field(self%h, 2, c_loc(self%gg%C)), "solid_earth_update")
      call check(self, vh_get_field(self%h, 4, c_loc(self%gg%s_rot)), "solid_earth_update")
      call check(self, vh_get_field(self%h, 5, c_loc(self%gg%h_ice)), "solid_earth_update")
      call check(self, vh_get_spectrum(self%h, 0, c_loc(self%sigma_lm)), "solid_earth_update")
      call sync_host_fields(self)   ! solid_earth_sync_host (src/vilma_coupling.f90:439-451)
      call check(self, vh_get_scalar(self%h, 0, v), "solid_earth_update"); self%time = v
      call check(self, vh_get_scalar(self%h, 1, v), "solid_earth_update"); self%bsl = v
      call check(self, vh_get_scalar(self%h, 2, v), "solid_earth_update"); self%worst_mass_resid = v
   end subroutine refresh

   !! solid_earth_init (src/vilma_coupling.f90:144-232); the caller sets
   !! self%par (e.g. by vilma_par_load) first. With a grid that is not the
   !! model's Gauss grid the reference fields come on the host grid and are
   !! mapped in (conservative; the ice with its mass kept), and rsl / z_bed are
   !! returned on it (bilinear), as the reference's host-grid coupling does.
   subroutine solid_earth_init(self, z_bed_eq, h_ice_eq, grid, h_ice_init)
      type(solid_earth), intent(inout), target :: self
      real(wp), intent(in), target :: z_bed_eq(:,:), h_ice_eq(:,:)
      type(grid_class), intent(in), optional :: grid
      real(wp), intent(in), target, optional :: h_ice_init(:,:)
      type(vh_params_c), target :: c
      type(c_ptr) :: hi0
      integer(c_int) :: nlm
      integer :: gnphi, gnlat
      real(wp), allocatable, target :: z(:,:), hi(:,:), h0(:,:)
      self%remap = .false.
      if (present(grid)) then
         call gauss_dims(self%par, gnphi, gnlat)
         self%remap = .not. grid_is_gauss(grid, gnphi, gnlat)
      end if
      if (self%remap) then
         if (size(z_bed_eq, 1) /= grid%G%nx .or. size(z_bed_eq, 2) /= grid%G%ny) &
            error stop 'solid_earth_init: z_bed_eq must be on the host grid (nx, ny)'
         call remap_init_dims(self%map, gnphi, gnlat, grid%G%x, grid%G%y)
         self%nphi = gnphi;  self%nlat = gnlat
         self%z_bed_eq = z_bed_eq
         allocate(z(gnphi, gnlat), hi(gnphi, gnlat))
         call map_in(self%map, z_bed_eq, z, conserve_mass=.false.)     ! bed: geometry
         call map_in(self%map, h_ice_eq, hi, conserve_mass=.true.)     ! ice: mass
      else
         self%nphi = size(z_bed_eq, 1); self%nlat = size(z_bed_eq, 2)
         z = z_bed_eq; hi = h_ice_eq
      end if
      c = to_c(self%par)
      hi0 = c_null_ptr
      if (present(h_ice_init)) then
         if (self%remap) then
            allocate(h0(gnphi, gnlat))
            call map_in(self%map, h_ice_init, h0, conserve_mass=.true.)
         else
            h0 = h_ice_init
         end if
         hi0 = c_loc(h0)
      end if
      self%h = c_null_ptr
      call check(self, vh_create(c, int(self%nphi, c_int), int(self%nlat, c_int), c_loc(z), c_loc(hi), hi0, &
                                 self%h), "solid_earth_init")
      call check(self, vh_get_dims(self%h, 3, nlm), "solid_earth_init")
      if (self%remap) then
         allocate(self%rsl(grid%G%nx, grid%G%ny), self%z_bed(grid%G%nx, grid%G%ny), self%sigma_lm(nlm))
      else
         allocate(self%rsl(self%nphi, self%nlat), self%z_bed(self%nphi, self%nlat), self%sigma_lm(nlm))
      end if
      allocate(self%gg%rsl(self%nphi, self%nlat), self%gg%z_bed(self%nphi, self%nlat), &
               self%gg%C(self%nphi, self%nlat), self%gg%C0(self%nphi, self%nlat), &
               self%gg%h_ice(self%nphi, self%nlat), self%gg%s_rot(self%nphi, self%nlat))
      call check(self, vh_get_field(self%h, 3, c_loc(self%gg%C0)), "solid_earth_init")
      self%h_ice_eq = h_ice_eq
      if (present(h_ice_init)) then
         self%h_ice = h_ice_init
      else
         self%h_ice = h_ice_eq
      end if
      if (.not. self%remap) self%z_bed_eq = z_bed_eq
      call refresh(self)
   end subroutine solid_earth_init

   !! solid_earth_update (src/vilma_coupling.f90:336-437)
   subroutine solid_earth_update(self, h_ice, dt_yr)
      type(solid_earth), intent(inout), target :: self
      real(wp), intent(in) :: h_ice(:,:)
      real(wp), intent(in) :: dt_yr
      real(wp), allocatable, target :: hi(:,:)
      if (.not. c_associated(self%h)) error stop 'solid_earth_update: model not initialised'
      if (self%remap) then   ! ice in conservatively, its mass kept (src/vilma_coupling.f90:355-359)
         allocate(hi(self%nphi, self%nlat))
         call map_in(self%map, h_ice, hi, conserve_mass=.true.)
      else
         if (size(h_ice, 1) /= self%nphi .or. size(h_ice, 2) /= self%nlat) &
            error stop 'solid_earth_update: h_ice must be on the model grid (nphi, nlat)'
         hi = h_ice
      end if
      call check(self, vh_update(self%h, c_loc(hi), real(dt_yr, c_double)), "solid_earth_update")
      self%h_ice = h_ice
      call refresh(self)
   end subroutine solid_earth_update

   !! solid_earth_spinup (src/vilma_coupling.f90:453-517): relax to isostatic
   !! equilibrium under the start-slice ice h_ice_lgm (host grid), holding the
   !! reference state as the datum, as se%par asks (pre_spinup_1d,
   !! equil_time_max, equil_rate_tol); a no-op when both are off.
   subroutine solid_earth_spinup(self, h_ice_lgm)
      type(solid_earth), intent(inout), target :: self
      real(wp), intent(in) :: h_ice_lgm(:,:)
      real(wp), allocatable, target :: hi(:,:)
      integer(c_int) :: pre1d
      if (.not. c_associated(self%h)) error stop 'solid_earth_spinup: model not initialised'
      if (.not. self%par%pre_spinup_1d .and. self%par%equil_time_max <= 0.0_wp) return
      if (self%remap) then
         allocate(hi(self%nphi, self%nlat))
         call map_in(self%map, h_ice_lgm, hi, conserve_mass=.true.)
      else
         if (size(h_ice_lgm, 1) /= self%nphi .or. size(h_ice_lgm, 2) /= self%nlat) &
            error stop 'solid_earth_spinup: h_ice_lgm must be on the model grid (nphi, nlat)'
         hi = h_ice_lgm
      end if
      pre1d = 0
      if (self%par%pre_spinup_1d) pre1d = 1
      call check(self, vh_spinup(self%h, c_loc(hi), real(self%par%equil_time_max / sec_per_year, c_double), &
                 real(self%par%equil_rate_tol, c_double), pre1d), "solid_earth_spinup")
      self%h_ice = h_ice_lgm
      call refresh(self)
   end subroutine solid_earth_spinup

   !! solid_earth_sync_host (src/vilma_coupling.f90:439-451): the host-grid
   !! rsl (bilinear from the Gauss grid, or a copy) and z_bed = z_bed_eq - rsl
   !! from the model's current state
   subroutine solid_earth_sync_host(self)
      type(solid_earth), intent(inout), target :: self
      if (.not. c_associated(self%h)) error stop 'solid_earth_sync_host: model not initialised'
      call check(self, vh_get_field(self%h, 0, c_loc(self%gg%rsl)), "solid_earth_sync_host")
      call sync_host_fields(self)
   end subroutine solid_earth_sync_host

   subroutine sync_host_fields(self)
      type(solid_earth), intent(inout) :: self
      if (self%remap) then
         call remap_to_ll(self%map, self%gg%rsl, self%rsl)
      else
         self%rsl = self%gg%rsl
      end if
      self%z_bed = self%z_bed_eq - self%rsl
   end subroutine sync_host_fields

   !! solid_earth_enable_visc_3d (src/vilma_coupling.f90:320-334) with the
   !! lateral field given in memory: log10-viscosity perturbation per element,
   !! (nphi, nlat, ne) (response_enable_lateral_visc, src/vilma_response.f90:994-1053)
   subroutine solid_earth_enable_visc_3d_field(self, pert_elem)
      type(solid_earth), intent(inout) :: self
      real(wp), intent(in), target, contiguous :: pert_elem(:,:,:)
      if (size(pert_elem, 1) /= self%nphi .or. size(pert_elem, 2) /= self%nlat .or. &
          size(pert_elem, 3) /= dim_of(self, 5)) &
         error stop 'enable_lateral_visc: pert_elem must be (nphi,nlat,ne)'
      call check(self, vh_enable_visc_3d(self%h, c_loc(pert_elem)), "solid_earth_enable_visc_3d")
   end subroutine solid_earth_enable_visc_3d_field

   !! the same from absolute nodal log10(eta), (nphi*nlat, nr)
   !! (response_enable_lateral_visc_from_nodes, src/vilma_response.f90:1120-1169)
   subroutine solid_earth_enable_visc_3d_nodes(self, visc_node)
      type(solid_earth), intent(inout) :: self
      real(wp), intent(in), target, contiguous :: visc_node(:,:)
      if (size(visc_node, 1) /= self%nphi * self%nlat .or. size(visc_node, 2) /= dim_of(self, 4)) &
         error stop 'enable_lateral_visc_from_nodes: visc_node must be (nphi*nlat, nr)'
      call check(self, vh_enable_visc_3d_nodes(self%h, c_loc(visc_node), self%par%visc3d_lid_depth, &
                 self%par%visc3d_lid_log10max, self%par%visc_log10_max), "solid_earth_enable_visc_3d")
   end subroutine solid_earth_enable_visc_3d_nodes

   !! The prognostic state vilma_restart_write / vilma_restart_read carry
   !! (src/vilma_io.f90:425-491), as arrays in the reference's layouts, for a
   !! host that keeps its own restart files: time (yr), tau(nlam, ne, nk, 6)
   !! = tau_a_re, tau_a_im, tau_b_re, tau_b_im, tau_c_re, tau_c_im, the rsl and
   !! ice of the Gauss grid (nphi, nlat), the polar motion m and both rotation
   !! channels' memory (nlam, ne, 6) (rotation_get_memory, src/vilma_rotation.f90:275-285).
   !! optional: the adaptive controller's next step dt_try [yr] and, for
   !! scheme "trap", the start-of-step load sigma_n(nlm) (zero and primed =
   !! .false. before the first trapezoidal step), src/vilma_io.f90:453-457,548-551
   subroutine solid_earth_get_restart_state(self, time, tau, rsl, h_ice, m, load_mem, tidal_mem, dt_try, &
                                            sigma_n, sigma_primed)
      type(solid_earth), intent(inout) :: self
      real(wp), intent(out) :: time
      real(wp), intent(out), target, contiguous :: tau(:,:,:,:), rsl(:,:), h_ice(:,:), load_mem(:,:,:), tidal_mem(:,:,:)
      complex(wp), intent(out) :: m
      real(wp), intent(out), optional :: dt_try
      complex(wp), intent(out), optional, target, contiguous :: sigma_n(:)
      logical, intent(out), optional :: sigma_primed
      real(wp), target :: mm(2)
      real(c_double) :: t, dtt
      integer(c_int) :: pr
      type(c_ptr) :: ps
      call check_restart_shapes(self, tau, rsl, h_ice, load_mem, tidal_mem)
      call check(self, vh_get_memory(self%h, c_loc(tau)), "solid_earth_get_restart_state")
      call check(self, vh_get_state(self%h, t, c_loc(rsl), c_loc(h_ice), c_loc(mm), c_loc(load_mem), &
                 c_loc(tidal_mem)), "solid_earth_get_restart_state")
      time = t
      m = cmplx(mm(1), mm(2), wp)
      ps = c_null_ptr
      if (present(sigma_n)) then
         if (size(sigma_n) /= size(self%sigma_lm)) error stop 'restart state: sigma_n must have nlm values'
         ps = c_loc(sigma_n)
      end if
      call check(self, vh_get_step_state(self%h, dtt, pr, ps), "solid_earth_get_restart_state")
      if (present(dt_try)) dt_try = dtt
      if (present(sigma_primed)) sigma_primed = pr /= 0
   end subroutine solid_earth_get_restart_state

   subroutine solid_earth_set_restart_state(self, time, tau, rsl, h_ice, m, load_mem, tidal_mem, dt_try, &
                                            sigma_n)
      type(solid_earth), intent(inout) :: self
      real(wp), intent(in) :: time
      real(wp), intent(in), target, contiguous :: tau(:,:,:,:), rsl(:,:), h_ice(:,:), load_mem(:,:,:), tidal_mem(:,:,:)
      complex(wp), intent(in) :: m
      real(wp), intent(in), optional :: dt_try
      complex(wp), intent(in), optional, target, contiguous :: sigma_n(:)
      real(wp), target :: mm(2)
      type(c_ptr) :: ps
      integer(c_int) :: pr
      real(c_double) :: dtt
      mm = [real(m, wp), aimag(m)]
      call check_restart_shapes(self, tau, rsl, h_ice, load_mem, tidal_mem)
      call check(self, vh_set_memory(self%h, c_loc(tau)), "solid_earth_set_restart_state")
      call check(self, vh_set_state(self%h, real(time, c_double), c_loc(rsl), c_loc(h_ice), c_loc(mm), &
                 c_loc(load_mem), c_loc(tidal_mem)), "solid_earth_set_restart_state")
      if (present(dt_try) .or. present(sigma_n)) then
         ! what is not given keeps its current value
         call check(self, vh_get_step_state(self%h, dtt, pr, c_null_ptr), "solid_earth_set_restart_state")
         if (present(dt_try)) dtt = dt_try
         ps = c_null_ptr
         if (present(sigma_n)) then
            if (size(sigma_n) /= size(self%sigma_lm)) error stop 'restart state: sigma_n must have nlm values'
            ps = c_loc(sigma_n)
            pr = 1
         else
            pr = 0
         end if
         call check(self, vh_set_step_state(self%h, dtt, pr, ps), "solid_earth_set_restart_state")
      end if
      call refresh(self)
   end subroutine solid_earth_set_restart_state

   !! read_ve_state's memory (src/vilma_io.f90:492-552): tau(nlam, ne, nk_file, 6)
   !! of a run at the same truncation, or at a lower one (cross-resolution:
   !! the shared low-degree block, higher degrees zero; the spatial fields and
   !! sigma_n are not carried over, so the host seeds with a zero-length update)
   subroutine solid_earth_set_restart_memory(self, tau)
      type(solid_earth), intent(inout) :: self
      real(wp), intent(in), target, contiguous :: tau(:,:,:,:)
      integer :: nlam, ne, nk
      nlam = dim_of(self, 8);  ne = dim_of(self, 5);  nk = dim_of(self, 6)
      if (size(tau, 1) /= nlam .or. size(tau, 2) /= ne .or. size(tau, 4) /= 6) &
         error stop 'restart memory: tau must be (nlam, ne, nk_file, 6)'
      if (size(tau, 3) == nk) then
         call check(self, vh_set_memory(self%h, c_loc(tau)), "vilma_restart_read")
      else
         call check(self, vh_set_memory_lowres(self%h, c_loc(tau), int(size(tau, 3), c_int)), "vilma_restart_read")
      end if
   end subroutine solid_earth_set_restart_memory

   !! a dimension of the model (vh_get_dims selector: 4 nr, 5 ne, 6 nk, 8 nlam)
   integer function dim_of(self, which) result(n)
      type(solid_earth), intent(inout) :: self
      integer, intent(in) :: which
      integer(c_int) :: v
      call check(self, vh_get_dims(self%h, int(which, c_int), v), "vilma")
      n = v
   end function dim_of

   !! the restart arrays must have the reference's shapes: the library reads
   !! and writes that many values through their addresses
   subroutine check_restart_shapes(self, tau, rsl, h_ice, load_mem, tidal_mem)
      type(solid_earth), intent(inout) :: self
      real(wp), intent(in) :: tau(:,:,:,:), rsl(:,:), h_ice(:,:), load_mem(:,:,:), tidal_mem(:,:,:)
      integer :: nlam, ne, nk
      nlam = dim_of(self, 8);  ne = dim_of(self, 5);  nk = dim_of(self, 6)
      if (size(tau, 1) /= nlam .or. size(tau, 2) /= ne .or. size(tau, 3) /= nk .or. size(tau, 4) /= 6) &
         error stop 'restart state: tau must be (nlam, ne, nk, 6)'
      if (size(rsl, 1) /= self%nphi .or. size(rsl, 2) /= self%nlat .or. &
          size(h_ice, 1) /= self%nphi .or. size(h_ice, 2) /= self%nlat) &
         error stop 'restart state: rsl and h_ice must be (nphi, nlat)'
      if (size(load_mem) /= 4 * ne * 6 .or. size(tidal_mem) /= 4 * ne * 6) &
         error stop 'restart state: rotation memory must be (4, ne, 6)'
   end subroutine check_restart_shapes

   !! vilma_par_print (src/vilma_params.f90:361): echo the record's active
   !! settings to stdout, or to `unit`
   subroutine vilma_par_print(p, unit)
      type(vilma_param_class), intent(in) :: p
      integer, intent(in), optional :: unit
      integer :: u, k
      u = 6
      if (present(unit)) u = unit
      write(u, '(a)') ' [vilma] configuration (MI355X library)'
      write(u, '(2a)') '   solver:   ', trim(p%solver)
      write(u, '(a,3(1x,i0))') '   lmax, nlat, nphi:', p%lmax, p%nlat, p%nphi
      write(u, '(2a)') '   earth:    ', trim(p%earth)
      if (trim(p%earth) == "custom") then
         do k = 1, p%n_layer
            write(u, '(a,i0,a,2es11.3,a,f8.1,a,2es11.3,a,i0,a,i0)') '     layer ', k, ' r ', p%r_bot(k), &
               p%r_top(k), ' rho ', p%rho(k), ' mu, eta ', p%mu(k), p%eta(k), ' rheology ', p%rheology(k), &
               ' elements ', p%n_elem_layer(k)
         end do
      end if
      write(u, '(4a,i0)') '   response: ', trim(p%earth_response), '   scheme: ', trim(p%scheme), p%max_couple_iter
      write(u, '(a,2(1x,i0),es10.2,2(1x,l1))') '   sle (n_outer, n_inner, tol, fixed_ocean, subgrid):', &
         p%sle_n_outer, p%sle_n_inner, p%sle_tol, p%sle_fixed_ocean, p%sle_subgrid
      write(u, '(a,4es11.3)') '   dt_init [s], rtol, atol, cfl:', p%dt_init, p%rtol, p%atol, p%cfl
      write(u, '(a,l1,2a)') '   rotation: ', p%rotation, '   deg1_frame: ', trim(p%deg1_frame)
      write(u, '(a,2es11.3,1x,l1)') '   spin-up (equil_time_max [s], equil_rate_tol, pre_spinup_1d):', &
         p%equil_time_max, p%equil_rate_tol, p%pre_spinup_1d
      if (p%l_visc_3d) then
         write(u, '(a,es10.2,a,2f7.2,a,l1)') '   lateral viscosity: visc3d_tol ', p%visc3d_tol, &
            '  clamp ', p%visc_log10_min, p%visc_log10_max, '  l_toroidal ', p%l_toroidal
         if (p%visc3d_lid_depth > 0.0_wp) write(u, '(a,f7.1,a,f6.2)') '     lid rule: above ', &
            p%visc3d_lid_depth * 1.0e-3_wp, ' km capped where log10(eta) > ', p%visc3d_lid_log10max
      end if
   end subroutine vilma_par_print

   !! nonzero status of a vh_remap_* call -> error stop with its message
   subroutine check_remap(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      type(c_ptr) :: m
      character(kind=c_char), pointer :: s(:)
      character(len=:), allocatable :: msg
      integer :: i, n
      if (rc == 0) return
      m = vh_last_error(c_null_ptr)
      n = int(c_strlen(m))
      call c_f_pointer(m, s, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = s(i)
      end do
      write(*, '(a)') trim(what) // ': ' // msg
      error stop 'vilma: ' // what // ' failed (message above)'
   end subroutine check_remap

   !! the model Gauss grid of a parameter record (build_sht, src/vilma_coupling.f90:644-662;
   !! NLAT_MIN = 32, src/vilma_sht.f90:72)
   subroutine gauss_dims(p, nphi, nlat)
      type(vilma_param_class), intent(in) :: p
      integer, intent(out) :: nphi, nlat
      nlat = p%nlat;  if (nlat <= 0) nlat = 2*p%lmax + 2
      nphi = p%nphi;  if (nphi <= 0) nphi = 4*p%lmax
      nlat = max(nlat, 32)
   end subroutine gauss_dims

   !! remap_init (src/vilma_remap.f90:56-99) for source cell-centre axes
   !! lon_src(:), lat_src(:) [degrees, ascending] against the Gauss grid of the
   !! model se (the reference passes its sht; no weight cache: the weights are
   !! rebuilt at each init, in milliseconds).
   subroutine remap_init(self, se, lon_src, lat_src)
      type(remap_ll_gauss), intent(out) :: self
      type(solid_earth), intent(in) :: se
      real(wp), intent(in) :: lon_src(:), lat_src(:)
      call remap_init_dims(self, se%nphi, se%nlat, lon_src, lat_src)
   end subroutine remap_init

   subroutine remap_init_dims(self, nphi, nlat, lon_src, lat_src)
      type(remap_ll_gauss), intent(out) :: self
      integer, intent(in) :: nphi, nlat
      real(wp), intent(in) :: lon_src(:), lat_src(:)
      real(wp), allocatable, target :: lon(:), lat(:)
      lon = lon_src;  lat = lat_src
      self%nlon = size(lon);  self%nlat_ll = size(lat)
      self%nphi = nphi;       self%nlat = nlat
      call check_remap(vh_remap_create(int(nphi, c_int), int(nlat, c_int), int(self%nlon, c_int), &
                       int(self%nlat_ll, c_int), c_loc(lon), c_loc(lat), self%m), "remap_init")
   end subroutine remap_init_dims

   !! remap_to_gauss (src/vilma_remap.f90:101-138): f_ll(nlon, nlat_ll) -> f_gauss(nphi, nlat)
   subroutine remap_to_gauss(self, se, f_ll, f_gauss, conserve_mass)
      type(remap_ll_gauss), intent(in) :: self
      type(solid_earth), intent(in) :: se
      real(wp), intent(in) :: f_ll(:,:)
      real(wp), intent(out) :: f_gauss(:,:)
      logical, optional, intent(in) :: conserve_mass
      if (se%nphi /= self%nphi .or. se%nlat /= self%nlat) error stop 'vilma_remap: map and model grids differ'
      call map_in(self, f_ll, f_gauss, conserve_mass)
   end subroutine remap_to_gauss

   subroutine map_in(self, f_ll, f_gauss, conserve_mass)
      type(remap_ll_gauss), intent(in) :: self
      real(wp), intent(in) :: f_ll(:,:)
      real(wp), intent(out) :: f_gauss(:,:)
      logical, optional, intent(in) :: conserve_mass
      real(wp), allocatable, target :: fi(:,:), fo(:,:)
      integer(c_int) :: cm
      if (size(f_ll,1) /= self%nlon .or. size(f_ll,2) /= self%nlat_ll) &
         error stop 'vilma_remap: source field shape /= map source grid'
      if (size(f_gauss,1) /= self%nphi .or. size(f_gauss,2) /= self%nlat) &
         error stop 'vilma_remap: target field shape /= SHTns grid'
      cm = 0
      if (present(conserve_mass)) then
         if (conserve_mass) cm = 1
      end if
      fi = f_ll
      allocate(fo(self%nphi, self%nlat))
      call check_remap(vh_remap_to_gauss(self%m, c_loc(fi), c_loc(fo), cm), "remap_to_gauss")
      f_gauss = fo
   end subroutine map_in

   !! remap_to_ll (src/vilma_remap.f90:140-164): f_gauss(nphi, nlat) -> f_ll(nlon, nlat_ll)
   subroutine remap_to_ll(self, f_gauss, f_ll)
      type(remap_ll_gauss), intent(in) :: self
      real(wp), intent(in) :: f_gauss(:,:)
      real(wp), intent(out) :: f_ll(:,:)
      real(wp), allocatable, target :: gi(:,:), fo(:,:)
      if (size(f_gauss,1) /= self%nphi .or. size(f_gauss,2) /= self%nlat) &
         error stop 'vilma_remap: source field shape /= SHTns grid'
      if (size(f_ll,1) /= self%nlon .or. size(f_ll,2) /= self%nlat_ll) &
         error stop 'vilma_remap: target field shape /= map target grid'
      gi = f_gauss
      allocate(fo(self%nlon, self%nlat_ll))
      call check_remap(vh_remap_to_ll(self%m, c_loc(gi), c_loc(fo)), "remap_to_ll")
      f_ll = fo
   end subroutine remap_to_ll

   !! release the map
   subroutine remap_free(self)
      type(remap_ll_gauss), intent(inout) :: self
      integer(c_int) :: rc
      if (c_associated(self%m)) rc = vh_remap_destroy(self%m)
      self%m = c_null_ptr
   end subroutine remap_free

   !! coords::grid_init for a lon-lat grid: the cell-centre axes in degrees
   subroutine grid_init(grid, name, mtype, units, lon180, x, y)
      type(grid_class), intent(out) :: grid
      character(len=*), intent(in), optional :: name, mtype, units
      logical, intent(in), optional :: lon180
      real(wp), intent(in) :: x(:), y(:)
      if (present(mtype)) then
         if (trim(mtype) /= "latlon") error stop 'grid_init: only mtype="latlon" grids are supported'
      end if
      if (present(units)) then
         if (trim(units) /= "degrees") error stop 'grid_init: axes must be in degrees'
      end if
      if (present(lon180)) continue
      if (present(name)) grid%name = name
      grid%G%nx = size(x);  grid%G%ny = size(y)
      grid%G%x = x;  grid%G%y = y
   end subroutine grid_init

   !! grid_is_gauss (src/vilma_coupling.f90:664-690): the host grid IS the
   !! model's Gauss grid (dimensions, and axes within 1e-6 degrees)
   logical function grid_is_gauss(grid, nphi, nlat) result(same)
      type(grid_class), intent(in) :: grid
      integer, intent(in) :: nphi, nlat
      real(wp), parameter :: TOL = 1.0e-6_wp, PI = 3.14159265358979323846_wp
      real(wp) :: x, p0, p1, p2, dp
      integer :: j, k, it
      same = .false.
      if (grid%G%nx /= nphi .or. grid%G%ny /= nlat) return
      do j = 1, nphi
         if (abs(grid%G%x(j) - 360.0_wp*real(j - 1, wp)/real(nphi, wp)) > TOL) return
      end do
      ! Gauss-Legendre nodes by Newton's method on P_nlat (ascending latitude)
      do j = 1, nlat
         x = -cos(PI*(real(j, wp) - 0.25_wp)/(real(nlat, wp) + 0.5_wp))
         do it = 1, 100
            p0 = 1.0_wp;  p1 = x
            do k = 2, nlat
               p2 = (real(2*k - 1, wp)*x*p1 - real(k - 1, wp)*p0)/real(k, wp)
               p0 = p1;  p1 = p2
            end do
            dp = real(nlat, wp)*(x*p1 - p0)/(x*x - 1.0_wp)
            if (abs(p1/dp) < 1.0e-15_wp) exit
            x = x - p1/dp
         end do
         if (abs(grid%G%y(j) - asin(x)*180.0_wp/PI) > TOL) return
      end do
      same = .true.
   end function grid_is_gauss

   !! solid_earth_finalize (src/vilma_coupling.f90:599-626)
   subroutine solid_earth_finalize(self)
      type(solid_earth), intent(inout) :: self
      integer(c_int) :: rc
      if (c_associated(self%h)) then
         rc = vh_destroy(self%h)
         self%h = c_null_ptr
      end if
      if (allocated(self%rsl)) deallocate(self%rsl, self%z_bed, self%sigma_lm)
      if (allocated(self%gg%rsl)) deallocate(self%gg%rsl, self%gg%z_bed, self%gg%C, self%gg%C0, &
                                             self%gg%h_ice, self%gg%s_rot)
      if (self%remap) call remap_free(self%map)
      if (allocated(self%z_bed_eq)) deallocate(self%z_bed_eq)
      self%remap = .false.
   end subroutine solid_earth_finalize

end module vilma
