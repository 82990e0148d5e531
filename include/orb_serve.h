/*
 * orb_serve.h -- serving ORB builds on one GPU with continuous batching.
 *
 * No counterpart in the reference (it runs one build per process,
 * src/orbit.cpp:22-289).  A server keeps a particle pool and a cell-heap pool
 * in HBM.  Requests are admitted first-in first-out when a build slot is
 * free and both pools can hold them (with zero capacities the particle pool
 * takes 40% of the HBM free at creation).  Each orb_server_step runs one
 * level of every admitted build in the same launches (the batched form's
 * kernels over the union of their cells), so builds that arrive while others
 * run join at the next level step.  A build's result depends only on its own
 * particles, so it equals the same build run alone (orb_build, batched form,
 * same level mode and semantics) bit for bit.
 */
#ifndef ORB_SERVE_H
#define ORB_SERVE_H

#include "orb.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orb_server orb_server;

typedef struct orb_server_config {
    int64_t particle_capacity;  /* particles resident at once; 0: from free HBM   */
    int64_t cell_capacity;      /* heap slots resident at once; 0: 2^22            */
    int32_t max_active;         /* builds admitted at once (1..4096)               */
    int32_t levels_mode;        /* ORB_LEVELS_* for every request                 */
    int32_t cmp;                /* ORB_CMP_*                                      */
    int32_t partition;          /* ORB_PART_*                                     */
} orb_server_config;

typedef struct orb_server_stats {
    int64_t submitted, admitted, completed;
    int64_t steps;              /* level steps launched                           */
    int64_t cells_stepped;      /* cells over all steps                           */
    int64_t max_batch;          /* most builds in one level step                  */
    int64_t pending, active, finishing;
    int64_t particle_capacity, particles_free;
    int64_t cell_capacity, cells_free;
} orb_server_stats;

typedef struct orb_request_info {
    int64_t n;
    int32_t d;
    int32_t error;              /* 0, ORB_ERR_UB (error_cell as orb_build reports
                                   it) or ORB_ERR_ARG (a NaN coordinate)         */
    int32_t error_cell;
    int32_t nlevels;
    int32_t total_passes;
    int32_t level_passes[64];
    int32_t pad;
    double submit_ms, admit_ms, done_ms;  /* steady-clock milliseconds          */
} orb_request_info;

int orb_server_create(orb_server **server, const orb_server_config *cfg);
int orb_server_destroy(orb_server *server);
/* Copies the particles (host SoA) into server-owned staging and queues the
 * build; ORB_ERR_BUSY if it can never fit the pools.                      */
int orb_server_submit(orb_server *server, const float *x, const float *y, const float *z,
                      int64_t n, int32_t d, uint64_t *request_id);
/* The same for particles already in device memory of the server's GPU:
 * they are copied into the pool at admission and written back, in their
 * final order, to the same arrays when the build ends.  The arrays stay
 * valid and untouched until orb_server_poll reports the request done.    */
int orb_server_submit_device(orb_server *server, float *x, float *y, float *z, int64_t n,
                             int32_t d, uint64_t *request_id);
/* Admits what fits, launches one level step of every active build, retires
 * finished builds.  Returns the number of requests not yet fetched-ready
 * (pending + active + finishing), or an error.                           */
int orb_server_step(orb_server *server);
/* 1 when the request's results are ready, 0 otherwise.                   */
int orb_server_poll(orb_server *server, uint64_t request_id);
/* Pops the id of a request that finished since the last call (completion
 * order); returns 1 with *request_id set, 0 when none is waiting.        */
int orb_server_next_done(orb_server *server, uint64_t *request_id);
/* Copies the results (any pointer may be NULL), releases the request.
 * ranges are relative to the request's particles ([0, n)); cap >= the
 * request's orb_heap_capacity(d).  For a device request x, y, z (host) get
 * a copy of the reordered device arrays.  Returns the build's status.     */
int orb_server_fetch(orb_server *server, uint64_t request_id, float *x, float *y, float *z,
                     uint32_t *ranges, orb_cell *heap, uint8_t *exists, int64_t cap,
                     orb_request_info *info);
int orb_server_get_stats(orb_server *server, orb_server_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* ORB_SERVE_H */
