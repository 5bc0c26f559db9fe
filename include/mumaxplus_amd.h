/*
 * mumaxplus_amd.h -- C ABI of the MI355X micromagnetic time-stepping core.
 *
 * Every entry point returns 0 on success or a non-zero error code
 * (MPA_ERR_*); mpa_last_error() gives the message of the last failure on the
 * calling thread.  No entry point exits the process.  Arrays are plain
 * pointers; field data is structure-of-arrays float32 with index
 * c*ncells + x + nx*(y + ny*z) (the reference's Field layout,
 * src/core/field.hpp:15-94 and src/bindings/fieldToArray.hpp:37-46).
 * Host arrays are copied; nothing keeps a pointer to caller memory.
 *
 * Each entry point names the reference interface it replaces
 * (dirkpetersen/mumax-plus, file:line).  The reference binds these classes to
 * Python through pybind11 (src/bindings/main.cpp:3-27); this library exposes
 * the same object model through handles.
 */
#ifndef MUMAXPLUS_AMD_H
#define MUMAXPLUS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPA_OK 0
#define MPA_ERR_INVALID 1 /* invalid argument (the reference throws std::invalid_argument) */
#define MPA_ERR_RUNTIME 2 /* runtime error (std::runtime_error)                          */
#define MPA_ERR_HIP 3     /* HIP / device failure (the reference exits, cudaerror.hpp:9-18) */
#define MPA_ERR_INDEX 4   /* index out of range                                          */

typedef struct mpa_world mpa_world;
typedef struct mpa_ferromagnet mpa_ferromagnet;
typedef struct mpa_antiferromagnet mpa_antiferromagnet;
typedef struct mpa_demag_plan mpa_demag_plan;

/* Message of the last error on this thread ("" if none). */
const char* mpa_last_error(void);
/* ABI version of this library. */
int mpa_abi_version(void);
/* 1 if a HIP device is usable, 0 otherwise (no other call needed first). */
int mpa_device_available(void);

/* ---------------------------------------------------------------- World -- */

/* World(cellsize, mastergrid, pbc_repetitions): wrap_world.cpp:40-45,
 * MumaxWorld ctor mumaxworld.cpp:28-31, World::World world.cpp:9-19. */
int mpa_world_create(const double cellsize[3], const int mastergrid_size[3], const int mastergrid_origin[3],
                     const int pbc_repetitions[3], mpa_world** out);
int mpa_world_destroy(mpa_world* w);
/* World.cellsize (wrap_world.cpp:46-47) */
int mpa_world_get_cellsize(const mpa_world* w, double out[3]);
/* World.bias_magnetic_field (wrap_world.cpp:48-49) */
int mpa_world_set_bias_magnetic_field(mpa_world* w, const double b[3]);
int mpa_world_get_bias_magnetic_field(const mpa_world* w, double out[3]);
/* World.add_ferromagnet(grid, geometry_array, regions_array, name): wrap_world.cpp:52-64,
 * MumaxWorld::addFerromagnet mumaxworld.cpp:57-80.  geometry: ncells bytes or NULL; regions:
 * ncells region indices (System, system.cpp:13-36) or NULL (every cell in region 0, no region
 * defined). */
int mpa_world_add_ferromagnet(mpa_world* w, const int size[3], const int origin[3], const uint8_t* geometry,
                              const unsigned int* regions, const char* name, mpa_ferromagnet** out);
/* World.ferromagnets (wrap_world.cpp:87-88) */
int mpa_world_num_ferromagnets(const mpa_world* w, int* out);
int mpa_world_ferromagnet_at(mpa_world* w, int index, mpa_ferromagnet** out);
/* World.get_ferromagnet(name) (wrap_world.cpp:79-81); *out = NULL when absent. */
int mpa_world_find_ferromagnet(mpa_world* w, const char* name, mpa_ferromagnet** out);
/* World.add_antiferromagnet(grid, geometry_array, regions_array, name): wrap_world.cpp:66-76,
 * MumaxWorld::addAntiferromagnet mumaxworld.cpp:82-103.  Its two sublattices are ferromagnet
 * handles (mpa_afm_sublattice) named "<name>:sublattice_1" / "<name>:sublattice_2". */
int mpa_world_add_antiferromagnet(mpa_world* w, const int size[3], const int origin[3], const uint8_t* geometry,
                                  const unsigned int* regions, const char* name, mpa_antiferromagnet** out);
/* World.antiferromagnets / get_antiferromagnet(name) (wrap_world.cpp:83-91); *out = NULL when absent. */
int mpa_world_num_antiferromagnets(const mpa_world* w, int* out);
int mpa_world_antiferromagnet_at(mpa_world* w, int index, mpa_antiferromagnet** out);
int mpa_world_find_antiferromagnet(mpa_world* w, const char* name, mpa_antiferromagnet** out);
/* World.set_pbc(mastergrid, reps) / set_pbc(reps) / unset_pbc / mastergrid / pbc_repetitions
 * (wrap_world.cpp:100-113, mumaxworld.cpp:288-342). */
int mpa_world_set_pbc(mpa_world* w, const int mastergrid_size[3], const int mastergrid_origin[3],
                      const int reps[3]);
int mpa_world_set_pbc_auto(mpa_world* w, const int reps[3]);
int mpa_world_unset_pbc(mpa_world* w);
int mpa_world_get_pbc(const mpa_world* w, int mastergrid_size[3], int mastergrid_origin[3], int reps[3]);
/* World.bounding_grid (wrap_world.cpp:97-99, mumaxworld.cpp:274-286) */
int mpa_world_bounding_grid(const mpa_world* w, int size[3], int origin[3]);
/* The HIP stream (hipStream_t) all work of this world is ordered on. */
int mpa_world_stream(mpa_world* w, void** stream);

/* ----------------------------------------------------------- Ferromagnet -- */

/* Magnet.name / system.grid / geometry (wrap_magnet.cpp:17-24, wrap_system.cpp:40-51) */
int mpa_fm_name(const mpa_ferromagnet* m, char* buf, int buflen);
int mpa_fm_grid(const mpa_ferromagnet* m, int size[3], int origin[3]);
int mpa_fm_geometry(const mpa_ferromagnet* m, uint8_t* out); /* ncells bytes */
int mpa_fm_cellsingeo(const mpa_ferromagnet* m, int* out);     /* System::cellsingeo system.cpp:80-83 */
/* Ferromagnet.enable_demag / enable_openbc (wrap_ferromagnet.cpp:33-38) and
 * Magnet.enable_as_stray_field_source / _destination / enable_elastodynamics (wrap_magnet.cpp:26-33,
 * Magnet::setEnableElastodynamics magnet.cpp:177-213). flag: "enable_demag", "enable_openbc",
 * "enable_as_stray_field_source", "enable_as_stray_field_destination", "enable_elastodynamics"
 * (not for a sublattice, nor while a rigid strain is set).  The stray fields of the other magnets
 * enter a destination's external field (zeeman.cu:70-88; a sublattice uses its antiferromagnet's
 * destination switch).  Also "enable_zhang_li_torque", "enable_slonczewski_torque",
 * "fixed_layer_on_top" (wrap_ferromagnet.cpp:35-39), and "noise_seed" (the 32-bit key of the
 * magnet's thermal field; the reference seeds cuRAND from the clock, ferromagnet.cpp:79-81, so it
 * has no such setting) with the read-only "noise_draws" (thermal draws taken by time steps;
 * MPA_ERR_INDEX once the 64-bit count exceeds an int: mpa_fm_get_noise_state reads it whole). */
int mpa_fm_set_flag(mpa_ferromagnet* m, const char* flag, int value);
int mpa_fm_get_flag(const mpa_ferromagnet* m, const char* flag, int* out);
/* The thermal-noise stream state of a ferromagnet or sublattice: the Philox key and the 64-bit
 * draw counters (draws taken by time steps, draws taken by thermal_noise evaluations).  No
 * counterpart in the reference (cuRAND seeded from the clock, ferromagnet.cpp:79-81, and no
 * checkpoints); a snapshot saves it so that a restored run continues the uninterrupted one. */
int mpa_fm_get_noise_state(const mpa_ferromagnet* m, uint32_t* seed, uint64_t* step_draws,
                           uint64_t* quantity_draws);
int mpa_fm_set_noise_state(mpa_ferromagnet* m, uint32_t seed, uint64_t step_draws, uint64_t quantity_draws);
/* Parameter / VectorParameter (wrap_parameter.cpp:13-107, parameter.cu:19-39).  key: msat, aex,
 * alpha, ku1, ku2, kc1, kc2, kc3, frozen_spins, anisU, anisC1, anisC2, bias_magnetic_field,
 * dmi_xxy, dmi_xyz, dmi_xxz, dmi_yxy, dmi_yyz, dmi_yxz, dmi_zxy, dmi_zyz, dmi_zxz, B1, B2,
 * temperature, xi, Lambda, free_layer_thickness, epsilon_prime, fixed_layer, pol, jcur
 * (wrap_ferromagnet.cpp:41-70); an independent ferromagnet also has the elastic parameters C11,
 * C12, C44, eta, rho, external_body_force, rigid_norm_strain, rigid_shear_strain
 * (wrap_magnet.cpp:35-43). */
int mpa_fm_param_info(const mpa_ferromagnet* m, const char* key, int* ncomp, int* is_uniform, char* name,
                      int namelen, char* unit, int unitlen);
int mpa_fm_param_set_uniform(mpa_ferromagnet* m, const char* key, const double* value /* ncomp */);
int mpa_fm_param_set_cells(mpa_ferromagnet* m, const char* key, const float* values /* ncomp*ncells */);
int mpa_fm_param_get(const mpa_ferromagnet* m, const char* key, float* out /* ncomp*ncells */);
int mpa_fm_param_get_uniform(const mpa_ferromagnet* m, const char* key, double* out /* ncomp */);
/* Time-dependent terms (Parameter.add_time_term / remove_time_terms / is_dynamic,
 * wrap_parameter.cpp:16-30 and 44-98; DynamicParameter, dynamic_parameter.hpp:10-68): the value at
 * time t is the static value + sum_k fn_k(t) (* mask_k per cell and component when given,
 * ncomp*ncells floats, copied).  fn writes ncomp values and returns 0, or non-zero to fail the
 * evaluation (the calling entry point then returns MPA_ERR_RUNTIME).  t is the time solver's time:
 * the stage time inside a Runge-Kutta step, World time otherwise.  mpa_fm_param_get returns the
 * value at that time; a masked term makes the parameter non-uniform (is_uniform 0). */
typedef int (*mpa_time_fn)(double t, void* ctx, double* out);
int mpa_fm_param_add_time_term(mpa_ferromagnet* m, const char* key, mpa_time_fn fn, void* ctx,
                               const float* mask /* ncomp*ncells or NULL */);
int mpa_fm_param_remove_time_terms(mpa_ferromagnet* m, const char* key);
int mpa_fm_param_is_dynamic(const mpa_ferromagnet* m, const char* key, int* out);
/* Electrical potential (PoissonSystem, poissonsystem.cu:134-307; wrap_poissonsystem.cpp;
 * LinSolver, linsolver.cpp and wrap_linsolver.cpp).  The parameters applied_potential (V, NaN
 * where none), conductivity (S/m) and amr_ratio are mpa_fm_param_* keys; the quantities
 * electrical_potential (1 component, solved on evaluation) and conductivity_tensor (6: xx yy zz xy
 * xz yz) go through mpa_fm_eval.  init builds the rows (the solution and the stepper are reset only
 * when the number of cells changes); solve = init + iterate, writing the potential (0 outside the
 * geometry); step / restart / residual (max norm of b - A x) / solver_solve (iterate, then copy
 * the FP64 solution when out is not NULL) act on the current system.  method: "jacobi"
 * (default), "conjugategradient", "minimalresidual", "steepestdescent".  Options: "max_iter"
 * (-1: no limit) and "tol" (default 1e-5). */
int mpa_fm_poisson_init(mpa_ferromagnet* m);
int mpa_fm_poisson_solve(mpa_ferromagnet* m, float* out /* ncells */);
int mpa_fm_poisson_step(mpa_ferromagnet* m);
int mpa_fm_poisson_restart(mpa_ferromagnet* m);
int mpa_fm_poisson_solver_solve(mpa_ferromagnet* m, double* out /* ncells or NULL */);
int mpa_fm_poisson_residual(mpa_ferromagnet* m, double* out);
int mpa_fm_poisson_set_method(mpa_ferromagnet* m, const char* name);
int mpa_fm_poisson_set_option(mpa_ferromagnet* m, const char* key, double value);
int mpa_fm_poisson_get_option(const mpa_ferromagnet* m, const char* key, double* out);
/* Parameter.set_in_region(idx, value) (wrap_parameter.cpp:37-38, 105-106; parameter.cu:41-47,
 * 144-150): the static value (ncomp doubles) in every cell of region idx; MPA_ERR_INVALID when idx
 * is not one of the magnet's regions (system.cpp:73-77). */
int mpa_fm_param_set_in_region(mpa_ferromagnet* m, const char* key, unsigned int region, const double* value);
/* System.regions (wrap_system.cpp:49-50): ncells indices (zeros without regions); the sorted
 * distinct indices (InterParameter.unique_regions, wrap_parameter.cpp:114-115): *count of them,
 * the first min(cap, *count) written to out (out may be NULL). */
int mpa_fm_regions(const mpa_ferromagnet* m, unsigned int* out /* ncells */);
int mpa_fm_unique_regions(const mpa_ferromagnet* m, unsigned int* out, int cap, int* count);
/* InterParameter (wrap_parameter.cpp:109-125, inter_parameter.cu): key inter_exchange or
 * scale_exchange (ferromagnet.cpp:26-27).  size = N (N - 1) / 2 values, N = largest region index
 * + 1 (at least 1), in the order i > j at i (i - 1) / 2 + j (inter_parameter.hpp:77-81).  set:
 * one value for every pair; set_between / get_between: the pair (i, j), both regions of the
 * magnet and i != j, else MPA_ERR_INVALID.  eval writes the size values. */
int mpa_fm_inter_info(const mpa_ferromagnet* m, const char* key, int* is_uniform, double* uniform_value,
                      long long* size, char* name, int namelen, char* unit, int unitlen);
int mpa_fm_inter_set(mpa_ferromagnet* m, const char* key, double value);
int mpa_fm_inter_set_between(mpa_ferromagnet* m, const char* key, unsigned int i, unsigned int j, double value);
int mpa_fm_inter_get_between(const mpa_ferromagnet* m, const char* key, unsigned int i, unsigned int j,
                             double* out);
int mpa_fm_inter_eval(const mpa_ferromagnet* m, const char* key, float* out /* size */);
/* Variable.set / get of the magnetisation (wrap_variable.cpp:8-16, wrap_field.cpp:19-68,
 * NormalizedVariable::set variable.cpp:76-79): count 3 (uniform) or 3*ncells. */
int mpa_fm_set_magnetization(mpa_ferromagnet* m, const float* values, long long count);
int mpa_fm_get_magnetization(mpa_ferromagnet* m, float* out);
/* FieldQuantity.eval / average of the quantity factories (wrap_ferromagnet.cpp:75-128,
 * wrap_magnet.cpp:61-75, wrap_fieldquantity.cpp:8-22).  q: magnetization, full_magnetization,
 * exchange_field, anisotropy_field, external_field, dmi_field, demag_field, magnetoelastic_field,
 * homogeneous_exchange_field, inhomogeneous_exchange_field, effective_field, torque, llg_torque,
 * damping_torque, spin_transfer_torque, thermal_noise (a new draw each evaluation),
 * magnetoelastic_force, exchange_energy_density, anisotropy_energy_density,
 * zeeman_energy_density, dmi_energy_density, demag_energy_density,
 * magnetoelastic_energy_density, homogeneous_exchange_energy_density,
 * inhomogeneous_exchange_energy_density, total_energy_density; for an independent ferromagnet
 * also the magnet quantities strain_tensor, stress_tensor (6 components xx yy zz xy xz yz),
 * elastic_force, effective_body_force, elastic_damping, elastic_acceleration, elastic_velocity,
 * elastic_displacement, kinetic_energy_density, elastic_energy_density, poynting_vector. */
int mpa_fm_quantity_ncomp(const mpa_ferromagnet* m, const char* q, int* ncomp);
int mpa_fm_eval(mpa_ferromagnet* m, const char* q, float* out /* ncomp*ncells */);
int mpa_fm_average(mpa_ferromagnet* m, const char* q, double* out /* ncomp */);
/* ScalarQuantity.eval: max_torque, exchange_energy, anisotropy_energy, zeeman_energy,
 * dmi_energy, demag_energy, magnetoelastic_energy, homogeneous_exchange_energy,
 * inhomogeneous_exchange_energy, total_energy, and for an independent ferromagnet
 * kinetic_energy, elastic_energy (wrap_scalarquantity.cpp, wrap_magnet.cpp:70-73). */
int mpa_fm_scalar(mpa_ferromagnet* m, const char* q, double* out);
/* StrayField(magnet, grid).eval with set_method("fft" | "brute") (wrap_strayfield.cpp:10-25,
 * strayfield.cpp:72-77): the field (tesla) of an independent ferromagnet on the grid (size,
 * origin), method 0 = FFT, 1 = direct sum; out holds 3 * ncells of the grid. */
int mpa_fm_stray_field(mpa_ferromagnet* m, const int size[3], const int origin[3], int method, float* out);
/* Ferromagnet.is_sublattice / hostMagnet (wrap_ferromagnet.cpp:32, ferromagnet.hpp hostMagnet):
 * *out = the antiferromagnet of a sublattice, NULL for an independent ferromagnet. */
int mpa_fm_host(mpa_ferromagnet* m, mpa_antiferromagnet** out);
/* Magnet.elastic_displacement / elastic_velocity Variable.set / get (wrap_magnet.cpp:22-25,
 * wrap_variable.cpp:8-16): which = "elastic_displacement" or "elastic_velocity"; count 3
 * (uniform) or 3*ncells; zero outside the geometry.  Needs enable_elastodynamics; not for a
 * sublattice. */
int mpa_fm_set_variable(mpa_ferromagnet* m, const char* which, const float* values, long long count);
int mpa_fm_get_variable(mpa_ferromagnet* m, const char* which, float* out /* 3*ncells */);
/* _demag_kernel(magnet) (wrap_magnet.cpp:56-59): reference kernel-grid layout (6, Kz, Ky, Kx).
 * Call with out = NULL to get ksize only. */
int mpa_fm_demag_kernel(mpa_ferromagnet* m, int ksize[3], float* out);

/* ------------------------------------------------------- Antiferromagnet -- */

/* Magnet.name / system.grid / geometry of an antiferromagnet (wrap_magnet.cpp:17-21). */
int mpa_afm_name(const mpa_antiferromagnet* a, char* buf, int buflen);
int mpa_afm_grid(const mpa_antiferromagnet* a, int size[3], int origin[3]);
int mpa_afm_geometry(const mpa_antiferromagnet* a, uint8_t* out);
int mpa_afm_cellsingeo(const mpa_antiferromagnet* a, int* out);
/* Antiferromagnet.sub1 / sub2 (wrap_antiferromagnet.cpp:19-20, antiferromagnet.cpp:33-39):
 * index 0 or 1.  The sublattice handles take every mpa_fm_* call. */
int mpa_afm_sublattice(mpa_antiferromagnet* a, int index, mpa_ferromagnet** out);
/* flag: "enable_as_stray_field_source", "enable_as_stray_field_destination",
 * "enable_elastodynamics" (wrap_magnet.cpp:26-33). */
int mpa_afm_set_flag(mpa_antiferromagnet* a, const char* flag, int value);
int mpa_afm_get_flag(const mpa_antiferromagnet* a, const char* flag, int* out);
/* Antiferromagnet parameters (wrap_antiferromagnet.cpp:24-30, antiferromagnet.cpp:15-25): afmex_cell,
 * afmex_nn, latcon, and the inter-sublattice DMI tensor dmi_xxy ... dmi_zxz; the elastic
 * parameters of the magnet (wrap_magnet.cpp:35-43): C11, C12, C44, eta, rho,
 * external_body_force, rigid_norm_strain, rigid_shear_strain. */
int mpa_afm_param_info(const mpa_antiferromagnet* a, const char* key, int* ncomp, int* is_uniform, char* name,
                       int namelen, char* unit, int unitlen);
int mpa_afm_param_set_uniform(mpa_antiferromagnet* a, const char* key, const double* value);
int mpa_afm_param_set_cells(mpa_antiferromagnet* a, const char* key, const float* values);
int mpa_afm_param_get(const mpa_antiferromagnet* a, const char* key, float* out);
int mpa_afm_param_get_uniform(const mpa_antiferromagnet* a, const char* key, double* out);
/* Time-dependent terms of an antiferromagnet parameter, as mpa_fm_param_add_time_term. */
int mpa_afm_param_add_time_term(mpa_antiferromagnet* a, const char* key, mpa_time_fn fn, void* ctx,
                                const float* mask);
int mpa_afm_param_remove_time_terms(mpa_antiferromagnet* a, const char* key);
int mpa_afm_param_is_dynamic(const mpa_antiferromagnet* a, const char* key, int* out);
/* Regions of an antiferromagnet (shared by its sublattices) and its inter-region parameters
 * inter_afmex_nn / scale_afmex_nn (antiferromagnet.cpp:22-23), as the mpa_fm_ ones. */
int mpa_afm_param_set_in_region(mpa_antiferromagnet* a, const char* key, unsigned int region,
                                const double* value);
int mpa_afm_regions(const mpa_antiferromagnet* a, unsigned int* out);
int mpa_afm_unique_regions(const mpa_antiferromagnet* a, unsigned int* out, int cap, int* count);
int mpa_afm_inter_info(const mpa_antiferromagnet* a, const char* key, int* is_uniform, double* uniform_value,
                       long long* size, char* name, int namelen, char* unit, int unitlen);
int mpa_afm_inter_set(mpa_antiferromagnet* a, const char* key, double value);
int mpa_afm_inter_set_between(mpa_antiferromagnet* a, const char* key, unsigned int i, unsigned int j,
                              double value);
int mpa_afm_inter_get_between(const mpa_antiferromagnet* a, const char* key, unsigned int i, unsigned int j,
                              double* out);
int mpa_afm_inter_eval(const mpa_antiferromagnet* a, const char* key, float* out);
/* Elastic displacement / velocity of the antiferromagnet, as mpa_fm_set_variable. */
int mpa_afm_set_variable(mpa_antiferromagnet* a, const char* which, const float* values, long long count);
int mpa_afm_get_variable(mpa_antiferromagnet* a, const char* which, float* out);
/* Quantities of the whole antiferromagnet: neel_vector, full_magnetization (neel.cu, fullmag.cu:13-18),
 * total_energy_density (energy.cu:70-76) and the magnet quantities listed at mpa_fm_eval. */
int mpa_afm_quantity_ncomp(const mpa_antiferromagnet* a, const char* q, int* ncomp);
int mpa_afm_eval(mpa_antiferromagnet* a, const char* q, float* out);
int mpa_afm_average(mpa_antiferromagnet* a, const char* q, double* out);
/* total_energy (energy.cu:78-92), kinetic_energy, elastic_energy (elasticenergies.cu:40-95). */
int mpa_afm_scalar(mpa_antiferromagnet* a, const char* q, double* out);
/* Stray field of the antiferromagnet (Ms1 m1 + Ms2 m2) on a grid, as mpa_fm_stray_field. */
int mpa_afm_stray_field(mpa_antiferromagnet* a, const int size[3], const int origin[3], int method, float* out);

/* ------------------------------------------------------------ TimeSolver -- */

/* TimeSolver properties (wrap_timesolver.cpp:15-44, timesolver.hpp:74-82).  key: time, timestep,
 * adaptive_timestep, max_error, headroom, lower_bound, upper_bound, sensible_factor,
 * sensible_timestep_default; read-only counters nsteps, nrejected, nrhs. */
int mpa_ts_get(const mpa_world* w, const char* key, double* out);
int mpa_ts_set(mpa_world* w, const char* key, double value);
/* TimeSolver.sensible_timestep (timesolver.cpp:37-47) */
int mpa_ts_sensible_timestep(mpa_world* w, double* out);
/* TimeSolver.set_method (timesolver.cpp:23-31; names of butchertableau.hpp:17-25) */
int mpa_ts_set_method(mpa_world* w, const char* name);
int mpa_ts_get_method(const mpa_world* w, char* buf, int buflen);
/* TimeSolver.step / steps / run (timesolver.cpp:77-111) */
int mpa_ts_step(mpa_world* w);
int mpa_ts_steps(mpa_world* w, long long n);
int mpa_ts_run(mpa_world* w, double duration);

/* ------------------------------------------------------ Relax / minimize -- */

/* World.relax(tol) / Ferromagnet.relax / Antiferromagnet.relax (wrap_world.cpp:95-96,
 * wrap_ferromagnet.cpp:72-73, wrap_antiferromagnet.cpp:31-32; Relaxer::exec relaxer.cpp:87-174):
 * damping-only Fehlberg dynamics while the total energy decreases, then with the error tolerance
 * lowered to tol while the torque decreases (or until the max torque is below the
 * RelaxTorqueThreshold where one is set).  The solver's method, tolerance, step mode, time and
 * step are restored afterwards. */
int mpa_world_relax(mpa_world* w, double tol);
int mpa_fm_relax(mpa_ferromagnet* m, double tol);
int mpa_afm_relax(mpa_antiferromagnet* a, double tol);
/* World.minimize / Ferromagnet.minimize / Antiferromagnet.minimize(tol, nsamples) (Minimizer,
 * minimizer.cu:8-160): steepest descent with Barzilai-Borwein steps until nsamples consecutive
 * maximum magnetisation changes are <= tol. */
int mpa_world_minimize(mpa_world* w, double tol, int nsamples);
int mpa_fm_minimize(mpa_ferromagnet* m, double tol, int nsamples);
int mpa_afm_minimize(mpa_antiferromagnet* a, double tol, int nsamples);
/* World.RelaxTorqueThreshold / Ferromagnet.RelaxTorqueThreshold (default -1;
 * wrap_world.cpp, wrap_ferromagnet.cpp:67). */
int mpa_world_set_relax_torque_threshold(mpa_world* w, double value);
int mpa_world_get_relax_torque_threshold(const mpa_world* w, double* out);
int mpa_fm_set_relax_torque_threshold(mpa_ferromagnet* m, double value);
int mpa_fm_get_relax_torque_threshold(const mpa_ferromagnet* m, double* out);

/* ------------------------------------------------------- Demag operator -- */

/* StrayFieldFFTExecutor (strayfieldfft.cu:149-246) + StrayFieldKernel (strayfieldkernel.cu:11-99)
 * as a stand-alone operator on device pointers. */
int mpa_demag_plan_create(const int size[3], const double cellsize[3], const int mastergrid_size[3],
                          const int pbc_repetitions[3], mpa_demag_plan** out);
int mpa_demag_plan_destroy(mpa_demag_plan* p);
/* h = -mu0 N * (Ms m): m, h device arrays of 3*ncells floats; msat_cells device array of ncells
 * floats or NULL (then msat_uniform is used); geometry device array of ncells bytes or NULL;
 * stream: hipStream_t or NULL. */
int mpa_demag_plan_exec(mpa_demag_plan* p, const float* m, const float* msat_cells, float msat_uniform,
                        const uint8_t* geometry, float* h, void* stream);
/* No counterpart in the reference: mpa_demag_plan_exec for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result; the FFT intermediates live in a static 64 MiB device scratch array
 * that each GPU holds until the process ends, instead of the plan's own buffers (which this call
 * never allocates).  A plan whose intermediates exceed 64 MiB (3 Nz Hx Ny plus, for Nz > 1,
 * 3 Hx Ky Nz complex FP32 values, Hx = Kx/2+1) is rejected with MPA_ERR_INVALID.  Calls share
 * one scratch per GPU: calls that overlap on one GPU (other streams or threads) race. */
int mpa_demag_plan_exec_noalloc(mpa_demag_plan* p, const float* m, const float* msat_cells, float msat_uniform,
                                const uint8_t* geometry, float* h, void* stream);
/* FFT lengths, algorithmic HBM bytes per exec and device bytes held by the plan. */
int mpa_demag_plan_info(const mpa_demag_plan* p, int fft_size[3], double* algorithmic_bytes,
                        double* device_bytes);

/* --------------------------------------------------------------- Timing -- */

/* Per-kernel timing of everything the world launches (no counterpart in the reference, which
 * has no timers).  While enabled, each kernel launch is bracketed by HIP events on the world
 * stream and its algorithmic HBM bytes are recorded.  The report gives, per kernel name (32
 * bytes each), the summed milliseconds, launch count and summed algorithmic bytes; *count
 * receives the number of kernel names (at most `cap` are written). */
int mpa_world_set_profiling(mpa_world* w, int enable);
/* Stepper switches (no counterpart in the reference).  "fuse_stages" (default 1): a world whose
 * only variable is one magnetisation on the uniform-parameter LLG path, stepped with a
 * first-same-as-last tableau, computes each stage state inside the field launch that produces
 * the last derivative it needs (no separate combination launch); so does a world whose only magnet
 * is one antiferromagnet on the sublattice-pair field kernel (any tableau; its displacement and
 * velocity, with elastodynamics, inside the acceleration launch); 0 runs the separate
 * combination, with the same arithmetic (the same values bit for bit).  "fuse_combine" (default 0): when the only variable is one
 * independent magnetisation with demag, each stage state is formed inside the demag's first FFT
 * pass, which reads the step start and the derivatives instead of the state (no separate
 * combination launch); 0 runs the separate combination.  "field_zchunk" (default 4): z planes per workgroup of the
 * uniform-parameter field kernel's LDS halo-tile variant (magnets with no geometry mask or periodic
 * wrap, at least 64 x 4 cells per plane); 0 selects the one-cell-per-thread variant; -1 lets the
 * tile-schedule autotuner pick 1, 2, 4 or 8 per grid size and kernel variant by timing each on the
 * first launch (mpa_field_tile_report).  Every chunk gives the same values bit for bit; the
 * one-cell-per-thread variant agrees with them to a few ulps.  "split_elastic" (default 0): 1 runs, with elastodynamics, the strain
 * and the acceleration of each right-hand side on another stream beside the demag passes; 0 runs
 * them on the world's stream (the same values).  "split_demag" (default 0): 1 computes the sublattice
 * fields without demag of an antiferromagnet with demag and uniform parameters beside the demag
 * passes and forms its torques after them (the same values bit for bit); 0 computes fields and
 * torques in one launch after the demag passes. */
int mpa_world_set_option(mpa_world* w, const char* key, int value);
int mpa_world_get_option(const mpa_world* w, const char* key, int* value);
/* The tile-schedule autotuner's record (process-wide; no counterpart in the reference), one line per
 * grid and field-kernel variant: "nx x ny x nz variant v | choice | 1 t1 us | 2 t2 us | 4 t4 us | 8 t8 us"
 * with "-" for a chunk not timed (not below nz).  At most cap bytes (NUL-terminated) are written to
 * text; *len = the full length.  mpa_field_tile_clear forgets every decision. */
int mpa_field_tile_report(int cap, char* text, int* len);
int mpa_field_tile_clear(void);
int mpa_world_profile_reset(mpa_world* w);
int mpa_world_profile_report(mpa_world* w, int cap, char* names, double* total_ms, long long* launches,
                             double* bytes, int* count);

/* ---------------------------------------------------------------- Batch -- */

/* Many independent small ferromagnets stepped together (no counterpart in the reference, where
 * each World is stepped on its own by TimeSolver::steps / run, timesolver.cpp:37-111; for callers
 * that run many small solves: parameter sweeps, ensembles, a request server).  A member is a
 * ferromagnet alone in its world, with uniform parameters and no elastodynamics; it keeps its own
 * step size, time, step controller and stopping rule.  Its state lives in the batch's packed
 * device arrays from mpa_batch_add until mpa_batch_remove, which copies it back into the magnet
 * and its world's time solver.  A member's world must outlive its membership.
 * capacity: packed cells (members occupy their cell count rounded up to 256); method: a tableau
 * whose last stage is the new state ("DormandPrince" or "BogackiShampine"). */
typedef struct mpa_batch mpa_batch;
int mpa_batch_create(long long capacity, const char* method, mpa_batch** out);
int mpa_batch_destroy(mpa_batch* b);
/* mode 0: `arg` accepted steps; 1: run for `arg` seconds (TimeSolver::run); 2: damping-only
 * dynamics until the max damping torque <= `arg` (1/s), or 200000 accepted steps when the torque
 * stalls above it (last_torque of mpa_batch_member then exceeds `arg`).  Adding 0x100 to the mode
 * keeps the world's current step instead of applying the sensible-step rule (a resumed member).  The member starts with its world's
 * solver settings (timestep, adaptive_timestep, max_error, headroom, bounds, sensible factor).
 * *id = -1 when its cells do not fit in the free capacity now. */
int mpa_batch_add(mpa_batch* b, mpa_ferromagnet* m, int mode, double arg, int* id);
int mpa_batch_remove(mpa_batch* b, int id);
/* n batch steps: one step attempt of every active member each. */
int mpa_batch_advance(mpa_batch* b, long long n);
/* Members whose stopping rule has been met (at most `cap` ids written; *count = all). */
int mpa_batch_finished(mpa_batch* b, int cap, int* ids, int* count);
int mpa_batch_member(mpa_batch* b, int id, double* time, double* dt, long long* steps, long long* rejected,
                     int* done, float* last_error, float* last_torque);
/* capacity, free cells, members, algorithmic HBM bytes and kernel launches of one batch step. */
int mpa_batch_info(mpa_batch* b, long long* capacity, long long* free_cells, int* members, double* step_bytes,
                   int* launches_per_step);
/* "graph": replay a captured HIP graph per step (default 1); "profiling": per-kernel HIP-event
 * timing (steps then run without the graph, one launch per pass);
 * "step_mode": 0 one launch per pass (default), 1 one persistent kernel for all n steps of an
 * advance (FFT lines up to 256 points), 2 whichever of these the autotuner measures faster for the
 * batch's size class; "persist_blocks_per_cu": resident blocks per compute unit of the persistent
 * kernel (0: as many as fit); "persist_max_grid": a cap on its grid (0: none); "persist_spin": the persistent kernel's grid-barrier poll limit
 * (0: its first barrier fails, for tests of the error path). */
int mpa_batch_set_option(mpa_batch* b, const char* key, int value);
/* The step path of the last advance (0 per-pass launches, 1 persistent kernel) and the grid of the
 * persistent kernel for the current members (0 when it cannot run them). */
int mpa_batch_step_path(mpa_batch* b, int* last_path, int* persist_grid);
/* The step-path autotuner's record (process-wide), one line per size class: the class, its choice
 * and the best measured ms per step of each candidate.  At most cap bytes (NUL-terminated) are
 * written to text; *len = the full length. */
int mpa_batch_tune_report(mpa_batch* b, int cap, char* text, int* len);
/* The autotuner's record written to a text file (one size class a line), merged from one (its
 * entries replace those of the same classes), or cleared; a process that loads a record skips the
 * timing of the classes it holds decisions for. */
int mpa_batch_tune_save(const char* path);
int mpa_batch_tune_load(const char* path);
int mpa_batch_tune_clear(void);
int mpa_batch_profile_report(mpa_batch* b, int cap, char* names, double* total_ms, long long* launches,
                             double* bytes, int* count);

/* ------------------------------------------------------- Measurement -- */
/* No counterpart in the reference: the measurement behind the FFT passes' choice of vector-ALU
 * butterflies over matrix-core products.  One radix-8 stage (8-point forward DFTs of `nlines`
 * lines, a multiple of 256, of 16 floats each: re/im interleaved) applied `rounds` (1..40) times
 * in LDS, by butterflies (variant 0) or as a 16 x 16 real matrix product on v_mfma_f32_16x16x4f32
 * (variant 1); in / out are host arrays of 16 nlines floats, *ms the kernel's time. */
int mpa_bench_dft8(int variant, int nlines, int rounds, const float* in, float* out, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* MUMAXPLUS_AMD_H */
