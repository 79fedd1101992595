/*
 * oracle.h -- CPU oracle for the NeuralSim simulation core.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only
 * as the checker.  The product library never links or calls it.
 *
 * Two halves:
 *   ref_*  FP64 restatement of the reference's own functions (file:line cited at each).
 *   mir_*  FP32 mirror of the device step: same op order, same integer fixed point, same
 *          Philox streams, so device results must equal it bit for bit.
 * Shared pieces (Philox4x32-10, float exp/log/cospi, Poisson tables) are restated here from
 * their published definitions, independently of the device code.
 */
#ifndef NSX_ORACLE_H
#define NSX_ORACLE_H
#include <stdint.h>
#include "../include/nsx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primitives ---------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_expf(float x);
float orc_logf(float x);
float orc_cospi(float x);
float orc_normal(uint32_t r0, uint32_t r1);       /* Box-Muller, float, own math        */
double orc_normal64(uint32_t r0, uint32_t r1);    /* same uniforms, libm double math     */
float orc_drive_normal(uint64_t seed, uint64_t i, uint32_t t); /* Gaussian drive of neuron i */
int orc_poisson_table(double lambda, uint32_t* thr, int cap); /* returns entries        */
/* FP32 mirror of one step's HH membrane integration of a single neuron (crossing mode): plain
 * Euler sub-steps, or the surrogate's sub-steps with rates frozen at the step-start V. */
void orc_hh_plain(const nsx_config* cfg, float* v, float* m, float* h, float* n, float I,
                  uint32_t substeps);
void orc_hh_frozen(const nsx_config* cfg, float* v, float* m, float* h, float* n, float I,
                   uint32_t substeps);

/* ---- FP64 reference restatement ------------------------------------------------------ */
int ref_lif_step(double* v, int32_t* refr, double resting, double threshold, double reset,
                 double tau_m, double refractory_period, double input_resistance, double dt,
                 double input_current);
int ref_izh_step(double* v, double* u, double a, double b, double c, double d, double dt,
                 double input_current);
double ref_hh_alpha_m(double v);
double ref_hh_beta_m(double v);
double ref_hh_alpha_h(double v);
double ref_hh_beta_h(double v);
double ref_hh_alpha_n(double v);
double ref_hh_beta_n(double v);
int ref_hh_step(double* v, double* m, double* h, double* n, double g_na, double g_k,
                double g_l, double e_na, double e_k, double e_l, double c_m, double dt,
                double input_current);
/* izhikevich.rs:20-52 with the V integration in `sub` sub-steps of dt/sub, u from the old v. */
int ref_izh_step_sub(double* v, double* u, double a, double b, double c, double d, double dt,
                     int sub, double input_current);
/* hodgkin_huxley.rs:46-86 gate and membrane update; documented = 1: the doc-comment equation
 * (hodgkin_huxley.rs:6-9), 0: the code's signs.  Returns the step-start V. */
double ref_hh_integrate(double* v, double* m, double* h, double* n, double g_na, double g_k,
                        double g_l, double e_na, double e_k, double e_l, double c_m, double dt,
                        double input_current, int documented);
/* One exact OU step of the drive current. */
double ref_ou_update(double x, double mean, double sigma, double tau, double dt, double z);
void ref_finalize(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                  uint64_t* row_ptr, uint32_t* col, uint64_t* perm);
double ref_mg_block(double mg_block_strength, double v);
double ref_conductance(double rise, double decay, double gmax, double t);
void ref_trace_on_pre(double* pre, double* last_pre, double time, double tau_plus);
void ref_trace_on_post(double* post, double* last_post, double time, double tau_minus);
void ref_trace_decay(double* pre, double* post, double dt, double tau_plus, double tau_minus);
double ref_stdp_weight_change(double dt, double a_plus, double a_minus, double tau_plus,
                              double tau_minus);
double ref_triplet_weight_change(double dt_pre_post, double dt_post_pre, double tau_plus,
                                 double tau_minus, double tau_y, double a_plus,
                                 double a_minus, double a_triplet);
double ref_rate_homeostasis(double mean_rate, double target_rate, int penalize_silence);
/* Literal per-synapse STDP loop of engine.rs:496-537 over a spike raster.
 * spikes_by_step: concatenated neuron ids, step_ptr[t]..step_ptr[t+1] for step t.
 * Synapses (src,tgt,w) are plastic iff plastic[k]. dt, times accumulate as net.time does. */
void ref_stdp_literal(uint64_t m, const uint32_t* src, const uint32_t* tgt, double* w,
                      const uint8_t* plastic, uint32_t n_steps, const uint64_t* step_ptr,
                      const uint32_t* spikes, uint64_t n_neurons, double dt, double a_plus,
                      double a_minus, double tau_plus, double tau_minus, double w_min,
                      double w_max);

/* FP64 network simulation with the reference's step functions and the device's schedule
 * (same CSR, delays, drives from the same Philox integers).  Used for tolerance checks. */
typedef struct ref_net ref_net;
ref_net* ref_net_create(const nsx_config* cfg, const uint64_t* row_ptr, const uint32_t* col,
                        const float* w, const uint8_t* delay);
void ref_net_destroy(ref_net* r);
/* run n steps, appending spikes (step,id) to the given buffers (cap); returns count */
uint64_t ref_net_run(ref_net* r, uint32_t n_steps, uint32_t* sp_step, uint32_t* sp_id,
                     uint64_t cap);
void ref_net_state(ref_net* r, double* v, double* u, double* m, double* h, double* n);
void ref_net_drive(ref_net* r, double* x);
void ref_net_weights(ref_net* r, double* w);

/* ---- FP32 mirror of the device engine ------------------------------------------------ */
typedef struct mir_net mir_net;
mir_net* mir_create(const nsx_config* cfg);
void mir_destroy(mir_net* r);
int mir_build_csr(mir_net* r, const uint32_t* src, const uint32_t* dst, const float* w,
                  const uint8_t* delay, uint64_t m);
int mir_generate_fixed_indegree(mir_net* r, uint32_t k_in, float w_exc_lo, float w_exc_hi,
                                float w_inh_lo, float w_inh_hi, uint32_t d_lo, uint32_t d_hi);
/* insertion-ordered synapse list of the generator (for finalize checks) */
void mir_generate_list(const nsx_config* cfg, uint32_t k_in, float w_exc_lo, float w_exc_hi,
                       float w_inh_lo, float w_inh_hi, uint32_t d_lo, uint32_t d_hi,
                       uint32_t* src, uint32_t* dst, float* w, uint8_t* delay);
uint64_t mir_num_synapses(mir_net* r);
void mir_read_csr(mir_net* r, uint64_t* row_ptr, uint32_t* col, float* w, uint8_t* delay);
/* run n steps; spikes appended (step,id) in step order then id order; returns count */
uint64_t mir_run(mir_net* r, uint32_t n_steps, uint32_t* sp_step, uint32_t* sp_id,
                 uint64_t cap, uint64_t* events_out);
void mir_read_state(mir_net* r, nsx_state_view* s);
void mir_write_state(mir_net* r, const nsx_state_view* s);
void mir_stimulate(mir_net* r, uint64_t neuron, float current);
uint64_t mir_step_index(mir_net* r);

#ifdef __cplusplus
}
#endif
#endif
