/* astro_canon.h -- physical constants of the per-step hot path.
 *
 * Every value is the double the reference simulator uses, taken from its
 * generated constant table (src/core/canon_generated.h:11-146) and its field
 * contract (contracts/fields_v1.h:130-152). The numbers must be identical for
 * results to be identical, so they are kept as the same IEEE doubles; the test
 * suite checks each one against the reference header compiled from source.
 *
 * Plain C so that the HIP kernels, the C++ host code and the CPU oracle all
 * read the same table. SI units throughout: m, kg, s, K, J, W.
 */
#ifndef ASTRO_CANON_H
#define ASTRO_CANON_H

/* ---- universal ---------------------------------------------------------- */
#define ASTRO_C_LIGHT        299792458.0
#define ASTRO_K_BOLTZ        1.3806490000000001e-23
#define ASTRO_G_ACCEL        9.8066499999999994
#define ASTRO_PI             3.14159265358979323846
#define ASTRO_TWO_PI         6.28318530717958647693

/* ---- cell ----------------------------------------------------------------- */
#define ASTRO_CELL_DIAMETER        1.0000000000000001e-05
#define ASTRO_CELL_RADIUS          5.0000000000000004e-06
#define ASTRO_CELL_VOLUME          5.2359877559829892e-16
#define ASTRO_CELL_CROSS_SECTION   7.8539816339744834e-11
#define ASTRO_CELL_SURFACE_AREA    3.1415926535897934e-10
#define ASTRO_CELL_MASS_DRY        2.0999999999999999e-14
#define ASTRO_CELL_TEMP_SETPOINT   369.565
#define ASTRO_CELL_ENERGY_MAX      1500000.0
#define ASTRO_CELL_ALBEDO_DEAD     0.84999999999999998
#define ASTRO_CELL_LETHAL_TEMP     573.14999999999998
#define ASTRO_PETROVA_MAX_POWER    0.050000000000000003
#define ASTRO_PETROVA_FLASH_POWER  3000000.0
/* the spin-drive flash audit's fixed-point scales (src/sim/flash.cu:26-27) */
#define ASTRO_FLASH_IMPULSE_SCALE  1.0e15
#define ASTRO_FLASH_ENERGY_SCALE   1.0e6

/* ---- biology -------------------------------------------------------------- */
#define ASTRO_CO2_MASS_PER_DIVISION   2.0999999999999999e-14
#define ASTRO_CO2_DIFFUSIVITY_WATER   1.92e-09
#define ASTRO_N2_DIFFUSIVITY          2.0e-09
#define ASTRO_LIFE_DIVISION_ENERGY_COST 0.0
#define ASTRO_LIFE_CO2_HALF_SATURATION 0.001
#define ASTRO_LIFE_CO2_UPTAKE_MAX     3.0381944444444441e-20

/* ---- water / medium ------------------------------------------------------- */
#define ASTRO_WATER_DENSITY        998.20000000000005
#define ASTRO_WATER_CONDUCTIVITY   0.59799999999999998
#define ASTRO_WATER_SPECIFIC_HEAT  4182.0
#define ASTRO_WATER_THERMAL_DIFFUSIVITY 1.4325163582348751e-07
#define ASTRO_VF_A                 2.4139999999999999e-05
#define ASTRO_VF_B                 570.58000000000004
#define ASTRO_VF_C                 140.0
#define ASTRO_AIR_CONVECTION_H     10.0
#define ASTRO_AMBIENT_TEMP_DEFAULT 293.14999999999998
#define ASTRO_CONDUCTION_COEFF     3.7573448136933928e-05

/* ---- mechanics -------------------------------------------------------------- */
#define ASTRO_CONTACT_STIFFNESS    1.1802e-05
#define ASTRO_WALL_STICKINESS      0.14999999999999999
#define ASTRO_WALL_STUCK_DRAG_MULT 20.0
#define ASTRO_WALL_RELEASE_FORCE   5.0000000000000002e-11
#define ASTRO_HASH_CELL_FACTOR     2.2000000000000002
#define ASTRO_DT_PHYSICS           0.001
#define ASTRO_CLOCK_PHYSICS_MIN    0.10000000000000001
#define ASTRO_CLOCK_PHYSICS_MAX    100.0
#define ASTRO_CLOCK_BIOLOGY_MIN    1.0
#define ASTRO_CLOCK_BIOLOGY_MAX    1000000.0
#define ASTRO_CLOCK_REALTIME_PHYSICS     1.0
#define ASTRO_CLOCK_REALTIME_BIOLOGY     1.0
#define ASTRO_CLOCK_MOTION_PHYSICS       10.0
#define ASTRO_CLOCK_MOTION_BIOLOGY       1.0
#define ASTRO_CLOCK_METABOLIC_PHYSICS    1.0
#define ASTRO_CLOCK_METABOLIC_BIOLOGY    10000.0
#define ASTRO_CLOCK_GENERATIONAL_PHYSICS 0.5
#define ASTRO_CLOCK_GENERATIONAL_BIOLOGY 1000000.0

/* ---- Taumoeba, the predator ---------------------------------------------------- */
#define ASTRO_TAU_DIAMETER          4.0000000000000003e-05
#define ASTRO_TAU_RADIUS            2.0000000000000002e-05
#define ASTRO_TAU_VOLUME            3.3510321638291131e-14
#define ASTRO_TAU_MASS              3.3450003059342215e-11
#define ASTRO_TAU_MASS_DRY          1e-13
#define ASTRO_TAU_CRAWL_SPEED       5.0000000000000004e-06
#define ASTRO_TAU_DRAG_20C          3.776489400816919e-07
#define ASTRO_TAU_CRAWL_THRUST      1.8882447004084599e-12
#define ASTRO_TAU_DIGEST_TIME       120.0
#define ASTRO_TAU_BIOMASS_YIELD     0.59999999999999998
#define ASTRO_TAU_DIVIDE_BIOMASS    2.0000000000000001e-13
#define ASTRO_TAU_N2_LETHAL_CONC    0.01
#define ASTRO_TAU_N2_LETHAL_TIME    2000.0
#define ASTRO_TAU_N2_HAZARD_RATE    0.050000000000000003
#define ASTRO_TAU_N2_TOLERANCE_INIT 0.0
#define ASTRO_TAU_N2_TOLERANCE_K    1.0
#define ASTRO_TAU_MUTATION_SIGMA    0.02

/* ---- taxis ------------------------------------------------------------------ */
#define ASTRO_TAXIS_MEMORY_TIME     0.10000000000000001
#define ASTRO_TAXIS_DARK_THRESHOLD  0.001
#define ASTRO_TAXIS_SEEK_FEED_BELOW 0.94999999999999996
#define ASTRO_TAXIS_SEEK_BREED_ABOVE 0.97999999999999998
#define ASTRO_TAXIS_TUMBLE_ANGLE_MEAN 1.1868238913561442
#define ASTRO_TAXIS_RUN_MAX         0.40000000000000002

/* ---- chamber and grids ------------------------------------------------------ */
#define ASTRO_CHAMBER_W            0.0040000000000000001
#define ASTRO_CHAMBER_H            0.0040000000000000001
#define ASTRO_CHAMBER_D            6.0000000000000002e-05
#define ASTRO_FIELD_N_TEMP         512
#define ASTRO_FIELD_N_CO2          256
#define ASTRO_FIELD_N_N2           128
#define ASTRO_FIELD_N_IRRAD        512
#define ASTRO_MAX_CELLS            2000000
#define ASTRO_DEFAULT_CELLS        25000
#define ASTRO_MAX_TAUMOEBA         65536
#define ASTRO_DEFAULT_TAUMOEBA     256

/* ---- fixed-point deposit scales (contracts/fields_v1.h:135-145) ------------ */
#define ASTRO_DEPOSIT_SCALE_TEMPERATURE 1.0e9
#define ASTRO_DEPOSIT_SCALE_CO2         1.0e15
#define ASTRO_DEPOSIT_SCALE_N2          1.0e15
#define ASTRO_DEPOSIT_SCALE_STATS       1.0e6
#define ASTRO_DEPOSIT_SCALE_IRRADIANCE  1.0e18
#define ASTRO_IRRADIANCE_NOMINAL_DIFFUSIVITY 1.0e-9
/* field sums in the telemetry reduction (src/sim/stats.cu:25) */
#define ASTRO_STATS_SCALE_FIELD         1.0e9

/* ---- cell flags (contracts/cell_store_v1.h:20-28) --------------------------- */
#define ASTRO_FLAG_OCCUPIED  (1u << 0)
#define ASTRO_FLAG_ALIVE     (1u << 1)
#define ASTRO_FLAG_AWAKE     (1u << 2)
#define ASTRO_FLAG_STUCK     (1u << 3)
#define ASTRO_FLAG_DIVIDING  (1u << 4)
#define ASTRO_FLAG_FLASHING  (1u << 5)

/* death causes (contracts/cell_store_v1.h:31-33) */
#define ASTRO_DEATH_NONE        0
#define ASTRO_DEATH_STARVED     1
#define ASTRO_DEATH_PREDATED    2
#define ASTRO_DEATH_OVERHEATED  3
#define ASTRO_DEATH_CULLED      4

#endif /* ASTRO_CANON_H */
