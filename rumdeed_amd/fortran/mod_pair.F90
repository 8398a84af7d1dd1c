! mod_pair -- particle bookkeeping of RUMDEED's hot path with the
! reference's interfaces (src/mod_pair.F90), on the device.
!
!   Add_Particle(par_pos, par_vel, par_species, step, emit, life, opt_sec)
!       src/mod_pair.F90:29-159: appends one particle with the Beeman history
!       seeded from the vacuum field; refuses it beyond the capacity
!       (nrPart_dropped), as MAX_PARTICLES does
!   Remove_Particles(step)  :352-562: order-preserving compaction of the
!       particles marked in the step's position update, with the life-time
!       histogram (record_lifetime, :1135-1160) copied into life_time
!
! `life` (particles_life) is read only by the ion collision model of the
! reference (src/mod_collisions.F90:123), which is not on the device path; it
! is accepted and not stored.
module mod_pair
  use, intrinsic :: iso_c_binding
  use mod_global
  use rumdeed_hip
  use rd_device
  implicit none
  private
  public :: Add_Particle, Remove_Particles

contains

  subroutine Add_Particle(par_pos, par_vel, par_species, step, emit, life, opt_sec)
    double precision, dimension(1:3), intent(in) :: par_pos, par_vel
    integer, intent(in) :: par_species, step, emit, life
    integer, intent(in), optional :: opt_sec
    double precision, target :: p(3), v(3)
    integer(c_int), target :: sp(1), sec(1)
    integer(c_int) :: n
    call RD_Require()
    p = par_pos
    v = par_vel
    sp(1) = int(par_species, c_int)
    sec(1) = 1
    if (present(opt_sec)) sec(1) = int(opt_sec, c_int)
    call rd_check(rd_add_particles(rd_ctx, 1_c_int, c_loc(p), c_loc(v), c_loc(sp), &
                  int(step, c_int), int(emit, c_int), c_loc(sec)), 'rd_add_particles')
    call rd_check(rd_particle_count(rd_ctx, n), 'rd_particle_count')
    if (n == nrPart) then
      nrPart_dropped = nrPart_dropped + 1
      return
    end if
    nrPart = n
    nrID = nrID + 1
    if (par_species == species_ion) then
      rd_ions = rd_ions + 1
      nrIon = nrIon + 1
    else
      nrElec = nrElec + 1
    end if
    if (allocated(particles_cur_pos)) then
      particles_cur_pos(:, n) = par_pos
      particles_cur_vel(:, n) = par_vel
      particles_species(n) = par_species
      particles_step(n) = step
      particles_emitter(n) = emit
      particles_section(n) = sec(1)
    end if
  end subroutine Add_Particle

  subroutine Remove_Particles(step)
    integer, intent(in) :: step
    type(rd_step_record) :: rec
    integer(c_int) :: hist(1000, 2)
    call RD_Require()
    call rd_check(rd_remove_particles(rd_ctx, int(step, c_int), rec), 'rd_remove_particles')
    call RD_Mirror(rec)
    call rd_check(rd_get_life_time(rd_ctx, hist), 'rd_get_life_time')
    life_time(1:MAX_LIFE_TIME, 1:2) = hist
    if (rd_sync_each_step) call RD_Sync_To_Host()
  end subroutine Remove_Particles
end module mod_pair
