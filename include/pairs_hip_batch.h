/*
 * pairs_hip_batch.h -- many independent LJ systems advanced together.
 *
 * The reference has no counterpart (it runs one simulation per process); this
 * is the build's batched ensemble: R independent periodic LJ replicas of
 * uneven sizes, each a full pairs_hip_ctx (include/pairs_hip.h, same modules,
 * same parity with the reference), whose non-rebuild timesteps run as ONE
 * persistent kernel launch for all replicas (fused force + final/initial
 * integrate + ghost-image refresh; as many workgroups as the device holds
 * resident, each walking a contiguous run of a table whose entry b is wave
 * b - base[r] of replica r), replayed from a captured hipGraph.  Rebuild
 * steps launch each rebuild kernel once over all replicas (a segment table
 * for the neighbour build).  The segment size (cells per workgroup of the
 * build, hence the neighbour-tile shape) is picked by an autotuner that
 * times the batched step over a set of candidates.
 *
 * All replicas share reneighbor_every, thermo_every, dt, cutoff, the list
 * form (full) and ntypes; they share the batch's stream.
 */
#ifndef PAIRS_HIP_BATCH_H
#define PAIRS_HIP_BATCH_H

#include "pairs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pairs_hip_batch pairs_hip_batch;

/* Create R replica contexts from R configurations (their stream fields are
   ignored: the batch's stream is used). */
int pairs_hip_batch_create(pairs_hip_batch **batch, int nreplicas, const pairs_hip_md_config *cfgs);
int pairs_hip_batch_destroy(pairs_hip_batch *batch);
const char *pairs_hip_batch_last_error(const pairs_hip_batch *batch);
/* Replica r's context, for set-up (pairs_hip_upload, pairs_hip_copper_fcc_lattice,
   pairs_hip_setup_cells) and downloads; owned by the batch. */
pairs_hip_ctx *pairs_hip_batch_replica(pairs_hip_batch *batch, int r);

/* Timesteps [ts_begin, ts_end) of the reference schedule for every replica,
   one batched launch per non-rebuild step.  Returns 0 or an error/overflow
   code of the first replica that failed. */
int pairs_hip_batch_run(pairs_hip_batch *batch, int ts_begin, int ts_end, int nsteps);
/* The same timesteps as a loop of per-replica pairs_hip_run calls (each
   replica with its own graphs): the comparison the batched form is timed
   against. */
int pairs_hip_batch_run_each(pairs_hip_batch *batch, int ts_begin, int ts_end, int nsteps);
/* Home particles per segment (the knob that sets cells per workgroup along z
   and thus the tile shape) for every replica; takes effect at the next rebuild. */
int pairs_hip_batch_set_home(pairs_hip_batch *batch, int home_per_segment);
/* At most `workgroups` workgroups in the persistent force launch (0, the
   default: as many as the device holds resident); each walks a contiguous run
   of the wave table.  Results do not depend on it. */
int pairs_hip_batch_set_workgroups(pairs_hip_batch *batch, int workgroups);
/* For each candidate home size: run the replicas to the next rebuild step,
   rebuild with that size and time `steps` batched steps (the replicas advance:
   trajectories do not depend on the segment size); keep the fastest.  Writes
   the chosen size, the per-candidate ms per step and the next timestep. */
int pairs_hip_batch_autotune(pairs_hip_batch *batch, int ts_begin, int steps, int nsteps, int ncandidates,
                             const int *candidates, double *ms_per_step, int *chosen, int *ts_next);
/* Last thermo row (T, P_kin, -, -, -) of each replica, [R][5]. */
int pairs_hip_batch_thermo(pairs_hip_batch *batch, double *out);

#ifdef __cplusplus
}
#endif
#endif
