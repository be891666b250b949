/*
 * oracle.h - CPU restatement of the reference's soft-step contact solve path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (twofold_amd/) may include, link or call this;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
 *
 * tfo_step() restates src/world.cpp:1591-1691 (step state, lists, SolverContext, the solve block)
 * followed by updateSleep (src/world.cpp:1174-1247), over the POD layouts of
 * include/twofold_solver.h. Each function in oracle.c cites the reference lines it follows.
 */
#ifndef TWOFOLD_ORACLE_H
#define TWOFOLD_ORACLE_H

#include "../include/twofold_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Integer results of a step. Arrays are caller-owned with room for n_contacts entries. */
typedef struct tfo_lists {
    uint32_t* active;
    uint32_t* fast;
    uint32_t* fast_solid;
    uint32_t* impact;
    uint32_t* by_color;         /* active contacts in row order: colours, then the rest */
    uint32_t* impact_by_color;  /* impact contacts in row order */
    uint8_t* colors;            /* per contact index; TFS_COLORS = none; 255 = not active */
    uint32_t n_active, n_fast, n_fast_solid, n_impact;
    uint32_t n_rest, n_impact_rest; /* trailing entries of by_color / impact_by_color */
    uint32_t n_colors;              /* highest colour used + 1 */
    uint32_t points_active, points_impact;
    /* reportTouches (src/world.cpp:1424-1468), when touch_events is not NULL: events as
     * (first contact of the pair << 1) | began, hits, and the touching contacts' count */
    uint32_t* touch_events;
    tfs_hit* hits;
    uint32_t n_touch_events, n_hits, touching;
    /* the contacts' reuse data, or NULL: a refresh leaves the contacts it collided again not reusable */
    tfs_contact_ext* ext;
} tfo_lists;

/* One step; returns 0, or TFS_E_CALLBACK when a refresh is due and refresh is NULL or fails. */
int tfo_step(tfs_body* bodies, uint32_t n_bodies, tfs_contact* contacts, uint32_t n_contacts,
             const tfs_settings* settings, float dt, tfs_refresh_fn refresh, void* user,
             tfo_lists* lists);

/* The contact stream (src/world.cpp:880-989): the contacts and reuse data updateContacts leaves, from
 * the old ones, the bodies, the candidate pairs and their manifolds. `out` and `out_ext` need room for
 * n_old + the manifolds' count entries. Returns the number of contacts. */
uint32_t tfo_stream_contacts(const tfs_body* bodies, const tfs_contact* old, const tfs_contact_ext* old_ext, uint32_t n_old,
                             const tfs_pair* pairs, uint32_t n_pairs, const tfs_manifold* manifolds,
                             const tfs_settings* settings, tfs_contact* out, tfs_contact_ext* out_ext);

/* The pieces, exposed for focused tests. */
void tfo_update_inverse_mass(tfs_body* b);
void tfo_make_soft(float hertz, float zeta, float h, float out[3]);
float tfo_maxf(float a, float b);
float tfo_minf(float a, float b);
/* The greedy colouring alone (lanes.cpp:356-374): colours of the listed contacts. */
void tfo_color(const tfs_body* bodies, uint32_t n_bodies, const tfs_contact* contacts,
               const uint32_t* list, uint32_t n, uint8_t* colors_by_contact);

#ifdef __cplusplus
}
#endif

#endif
