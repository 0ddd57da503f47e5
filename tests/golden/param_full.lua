-- Parameter file of a small plain COLA run: every key the driver reads,
-- with comment lines, trailing comments and options switched off by a
-- leading "--".
nc        = 48           -- particles per dimension
boxsize   = 75.0         -- Mpc/h

random_seed   = 2024
nrealization  = 2        -- seeds 2024 and 2025

ntimestep = 12
a_final   = 1.0
output_redshifts = {1.5, 0.5, 0.0}   -- snapshot times, highest z first

omega_m = 0.3
h       = 0.7
sigma8  = 0.8

pm_nc_factor    = 2      -- PM grid has pm_nc_factor * nc cells per side
np_alloc_factor = 1.5    -- particle buffer headroom
loglevel        = 1

powerspectrum = "inputs/pk_eh98.txt"   -- table of k [h/Mpc] and P(k)

--  Outputs below may be disabled by commenting the name out.

-- halo catalogue from friends-of-friends
fof            = "halos"
linking_factor = 0.18

-- full particle snapshot (GADGET-1)
snapshot = "snap"

-- random subsample of the particles
--subsample      = "subset"
subsample_factor = 0.05

-- CIC density grid
--    coarse_grid = "density"
coarse_grid_nc   = 24

-- particles at the first step, before any COLA step
-- initial = "ic"

write_longid = true      -- 8-byte particle IDs
