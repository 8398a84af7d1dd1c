! mod_global -- the module variables that RUMDEED's per-step hot path reads
! and writes, under the names, types and shapes of the reference's own
! module (src/mod_global.F90:26-506).  The drop-in modules of this directory
! (mod_verlet, mod_pair, mod_field_emission_v2, mod_field_thermo_emission,
! mod_emission_tip) use only names declared here, each declared as the
! reference declares it, so a program may build them against this module or
! against the reference's mod_global.F90 unchanged (the test suite compiles
! them against both).
!
! Only the declarations live here: the input namelist, the output units and
! the other physics models of the reference are outside the hot path.
module mod_global
  implicit none

  ! Physical constants (src/mod_global.F90:26-76)
  double precision, parameter :: pi = 3.141592653589793238462643383279502884197169399375105820974944592307816406286d0
  double precision, parameter :: q_0 = 1.602176634d-19       ! elementary charge (C)
  double precision, parameter :: m_0 = 9.1093837015d-31      ! electron mass (kg)
  double precision, parameter :: length_scale = 1.0d-9       ! 1 nm
  double precision, parameter :: time_scale = 1.0d-12        ! 1 ps
  double precision, parameter :: T_ntp = 293.15d0            ! K

  ! Sizes and species (src/mod_global.F90:98-108, :280)
  integer, parameter :: MAX_PARTICLES = 5000000
  integer, parameter :: MAX_EMITTERS = 1
  integer, parameter :: species_elec = 1
  integer, parameter :: species_ion = 2
  integer, parameter :: species_atom = 3
  integer, parameter :: nrSpecies = 3
  integer, parameter :: MAX_LIFE_TIME = 1000

  ! Particle arrays, (1:3, 1:MAX_PARTICLES) and (1:MAX_PARTICLES) (:128-152)
  double precision, dimension(:, :), allocatable :: particles_cur_pos
  double precision, dimension(:, :), allocatable :: particles_prev_pos
  double precision, dimension(:, :), allocatable :: particles_cur_vel
  double precision, dimension(:, :), allocatable :: particles_cur_accel
  double precision, dimension(:, :), allocatable :: particles_prev_accel
  double precision, dimension(:, :), allocatable :: particles_prev2_accel
  double precision, dimension(:), allocatable :: particles_charge
  integer, dimension(:), allocatable :: particles_species
  double precision, dimension(:), allocatable :: particles_mass
  integer, dimension(:), allocatable :: particles_step
  integer, dimension(:), allocatable :: particles_emitter
  integer, dimension(:), allocatable :: particles_section
  integer, dimension(:), allocatable :: particles_id

  ! Emitters (:182-187)
  double precision, dimension(:, :), allocatable :: emitters_pos
  double precision, dimension(:, :), allocatable :: emitters_dim
  integer, dimension(:), allocatable :: emitters_delay

  ! Input variables (:192-212, :313-314)
  double precision :: V_s
  double precision :: d
  double precision, dimension(1:3) :: box_dim
  double precision :: time_step
  double precision :: time_step2
  integer :: steps
  double precision :: T_temp = T_ntp
  logical :: image_charge = .true.
  integer :: N_ic_max = 0
  ! Two-time-step mode (:201, :363; set from atom_time_interval, main.F90:391-395)
  integer :: atom_time_interval = 0
  logical :: two_time_step = .false.

  ! Counters (:226-257)
  integer :: nrPart
  integer :: nrElec
  integer :: nrIon
  integer :: nrEmit
  integer :: nrID
  integer :: nrPart_remove
  integer :: nrElec_remove
  integer :: nrPart_remove_top
  integer :: nrPart_remove_bot
  integer :: nrElec_remove_top
  integer :: nrPart_dropped

  ! Ramo current per species, time, life-time histogram (:271-281)
  double precision, dimension(:), allocatable :: ramo_current
  double precision :: cur_time
  integer, dimension(:, :), allocatable :: life_time

  ! The emission plug-in (:445-506)
  interface
    subroutine Do_Emission(step)
      integer, intent(in) :: step
    end subroutine Do_Emission
  end interface
  procedure(Do_Emission), pointer :: ptr_Do_Emission => null()
end module mod_global
