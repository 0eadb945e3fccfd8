/*
 * hdm_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU FP64 restatement of HiDEM's per-timestep hot path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * and only as the checker / the timed CPU baseline.  See hdm_oracle.h for the
 * list of checks that pin each function.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, no fast-math).
 */
#include "hdm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int hdmo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Neighbour grid                                                            */
/* ------------------------------------------------------------------------ */

/* NeighbourGrid::getGridColor (NeighbourGrid.cpp:24-28). */
static int64_t grid_color(const hdmo_grid *g, int x, int y, int z)
{
    return (int64_t)x + (int64_t)y * (int64_t)g->X + (int64_t)z * (int64_t)g->X * (int64_t)g->Y;
}

static const int64_t *g_sort_keys;
static int cmp_colour_index(const void *pa, const void *pb)
{
    int32_t a = *(const int32_t *)pa, b = *(const int32_t *)pb;
    /* comparator of NeighbourGrid::generateIterationList (NeighbourGrid.cpp:172-179) */
    if (g_sort_keys[a] == g_sort_keys[b]) return (a < b) ? -1 : (a > b);
    return (g_sort_keys[a] < g_sort_keys[b]) ? -1 : 1;
}

/* NeighbourGrid::generate (NeighbourGrid.cpp:48-70) = generateNeighbourGrid
 * (:72-114) + colourParticles (:116-161) + generateIterationList (:163-184).
 * The colour -> bucket hash (OffloadHashMap) is replaced by a sorted array of
 * occupied colours searched by bisection; lookups give the same buckets. */
int hdmo_grid_build(const hdmo_particles *p, double searchDistance, hdmo_grid *g)
{
    int64_t n = p->n;
    memset(g, 0, sizeof(*g));
    if (n <= 1) return -1; /* ParticleData::getLimits exits on empty/degenerate data (ParticleData.h:52-56) */

    double xMin = p->x[0], xMax = p->x[0], yMin = p->y[0], yMax = p->y[0], zMin = p->z[0], zMax = p->z[0];
    for (int64_t i = 1; i < n; i++) {
        if (p->x[i] < xMin) xMin = p->x[i];
        if (p->x[i] > xMax) xMax = p->x[i];
        if (p->y[i] < yMin) yMin = p->y[i];
        if (p->y[i] > yMax) yMax = p->y[i];
        if (p->z[i] < zMin) zMin = p->z[i];
        if (p->z[i] > zMax) zMax = p->z[i];
    }
    g->spacing = searchDistance;
    g->xMin = xMin; g->yMin = yMin; g->zMin = zMin;
    g->xMax = xMax; g->yMax = yMax; g->zMax = zMax;
    /* getGridSize (NeighbourGrid.cpp:40-46) */
    g->X = 1 + (int)floor((xMax - xMin) / searchDistance);
    g->Y = 1 + (int)floor((yMax - yMin) / searchDistance);
    g->Z = 1 + (int)floor((zMax - zMin) / searchDistance);
    g->totalCells = (int64_t)g->X * g->Y * g->Z;

    int i = 0;
    for (int zD = -1; zD < 2; zD++)
        for (int yD = -1; yD < 2; yD++)
            for (int xD = -1; xD < 2; xD++)
                g->dirOffsets[i++] = grid_color(g, xD, yD, zD);  /* NeighbourGrid.cpp:100-108 */

    g->colours = (int64_t *)malloc(sizeof(int64_t) * n);
    g->order = (int32_t *)malloc(sizeof(int32_t) * n);
    if (!g->colours || !g->order) return -2;

#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
        /* getGridId (NeighbourGrid.cpp:30-33), colour stored through an int (:126) */
        int ix = (int)floor((p->x[k] - xMin) / searchDistance);
        int iy = (int)floor((p->y[k] - yMin) / searchDistance);
        int iz = (int)floor((p->z[k] - zMin) / searchDistance);
        int colour = (int)grid_color(g, ix, iy, iz);
        g->colours[k] = colour;
        g->order[k] = (int32_t)k;
    }
    g_sort_keys = g->colours;
    qsort(g->order, (size_t)n, sizeof(int32_t), cmp_colour_index);

    /* sort + unique colours -> buckets, sizes, exclusive-scan starts (:135-157) */
    int64_t nOcc = 0;
    for (int64_t k = 0; k < n; k++)
        if (k == 0 || g->colours[g->order[k]] != g->colours[g->order[k - 1]]) nOcc++;
    g->nOcc = nOcc;
    g->occColour = (int64_t *)malloc(sizeof(int64_t) * (nOcc ? nOcc : 1));
    g->bucketStart = (int32_t *)malloc(sizeof(int32_t) * (nOcc ? nOcc : 1));
    g->bucketSize = (int32_t *)malloc(sizeof(int32_t) * (nOcc ? nOcc : 1));
    int64_t b = -1;
    for (int64_t k = 0; k < n; k++) {
        int64_t c = g->colours[g->order[k]];
        if (k == 0 || c != g->colours[g->order[k - 1]]) {
            b++;
            g->occColour[b] = c;
            g->bucketStart[b] = (int32_t)k;
            g->bucketSize[b] = 0;
        }
        g->bucketSize[b]++;
    }
    return 0;
}

/* ---- redistribution order ---------------------------------------------------- */
typedef struct { int32_t c[4]; int64_t i; } reorder_key;

static int cmp_reorder_key(const void *pa, const void *pb)
{
    const reorder_key *a = (const reorder_key *)pa, *b = (const reorder_key *)pb;
    for (int k = 3; k >= 0; k--)   /* spacing 8s first (grids[3]) */
        if (a->c[k] != b->c[k]) return a->c[k] < b->c[k] ? -1 : 1;
    return (a->i < b->i) ? -1 : (a->i > b->i);
}

typedef struct { int32_t a, b; int64_t i; } link_key;

static int cmp_link_key(const void *pa, const void *pb)
{
    const link_key *x = (const link_key *)pa, *y = (const link_key *)pb;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->b != y->b) return x->b < y->b ? -1 : 1;
    return (x->i < y->i) ? -1 : (x->i > y->i);
}

static void permute_d(double *v, const int64_t *order, int64_t n, double *tmp)
{
    for (int64_t i = 0; i < n; i++) tmp[i] = v[order[i]];
    memcpy(v, tmp, sizeof(double) * (size_t)n);
}

int hdmo_reorder(hdmo_particles *p, hdmo_links *l, double searchDistance, int64_t *newOrder, int64_t *linkOrder)
{
    const int64_t n = p->n;
    if (n <= 1) return -1;
    double xMin = p->x[0], xMax = p->x[0], yMin = p->y[0], yMax = p->y[0], zMin = p->z[0], zMax = p->z[0];
    for (int64_t i = 1; i < n; i++) {
        if (p->x[i] < xMin) xMin = p->x[i];
        if (p->x[i] > xMax) xMax = p->x[i];
        if (p->y[i] < yMin) yMin = p->y[i];
        if (p->y[i] > yMax) yMax = p->y[i];
        if (p->z[i] < zMin) zMin = p->z[i];
        if (p->z[i] > zMax) zMax = p->z[i];
    }
    reorder_key *keys = (reorder_key *)malloc(sizeof(reorder_key) * (size_t)n);
    const int64_t L = l ? l->n : 0;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)((2 * L > n ? 2 * L : n) + 1));
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *inv = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!keys || !tmp || !order || !inv) {
        free(keys); free(tmp); free(order); free(inv);
        return -2;
    }
    for (int k = 0; k < 4; k++) {
        const double s = searchDistance * (double)(1 << k);
        hdmo_grid g;
        memset(&g, 0, sizeof(g));
        g.X = 1 + (int)floor((xMax - xMin) / s);
        g.Y = 1 + (int)floor((yMax - yMin) / s);
        for (int64_t i = 0; i < n; i++) {
            int ix = (int)floor((p->x[i] - xMin) / s);
            int iy = (int)floor((p->y[i] - yMin) / s);
            int iz = (int)floor((p->z[i] - zMin) / s);
            keys[i].c[k] = (int32_t)grid_color(&g, ix, iy, iz);
            keys[i].i = i;
        }
    }
    qsort(keys, (size_t)n, sizeof(reorder_key), cmp_reorder_key);
    for (int64_t i = 0; i < n; i++) {
        order[i] = keys[i].i;
        inv[keys[i].i] = i;
    }
    free(keys);
    double *fields[] = {p->x, p->y, p->z, p->tx, p->ty, p->tz, p->dx, p->dy, p->dz, p->du, p->dv, p->dw,
                        p->radius, p->fx, p->fy, p->fz, p->fu, p->fv, p->fw};
    for (size_t f = 0; f < sizeof(fields) / sizeof(fields[0]); f++) permute_d(fields[f], order, n, tmp);
    if (p->type) {
        uint8_t *t = (uint8_t *)tmp;
        for (int64_t i = 0; i < n; i++) t[i] = p->type[order[i]];
        memcpy(p->type, t, (size_t)n);
    }
    if (L > 0) {
        link_key *lk = (link_key *)malloc(sizeof(link_key) * (size_t)L);
        int64_t *lo = (int64_t *)malloc(sizeof(int64_t) * (size_t)L);
        int64_t *so = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * L));
        if (!lk || !lo || !so) {
            free(lk); free(lo); free(so); free(tmp); free(order); free(inv);
            return -2;
        }
        for (int64_t j = 0; j < L; j++) {
            lk[j].a = l->a[j];
            lk[j].b = l->b[j];
            lk[j].i = j;
        }
        qsort(lk, (size_t)L, sizeof(link_key), cmp_link_key);
        for (int64_t j = 0; j < L; j++) {
            lo[j] = lk[j].i;
            so[2 * j] = 2 * lk[j].i;
            so[2 * j + 1] = 2 * lk[j].i + 1;
        }
        free(lk);
        int32_t *ia = (int32_t *)tmp;
        for (int64_t j = 0; j < L; j++) ia[j] = (int32_t)inv[l->a[lo[j]]];
        int32_t *ib = (int32_t *)malloc(sizeof(int32_t) * (size_t)L);
        uint8_t *st = (uint8_t *)malloc((size_t)L);
        if (!ib || !st) {
            free(ib); free(st); free(lo); free(so); free(tmp); free(order); free(inv);
            return -2;
        }
        for (int64_t j = 0; j < L; j++) {
            ib[j] = (int32_t)inv[l->b[lo[j]]];
            st[j] = l->state[lo[j]];
        }
        memcpy(l->a, ia, sizeof(int32_t) * (size_t)L);
        memcpy(l->b, ib, sizeof(int32_t) * (size_t)L);
        memcpy(l->state, st, (size_t)L);
        free(ib);
        free(st);
        permute_d(l->initialLen, lo, L, tmp);
        if (l->maxLoad) permute_d(l->maxLoad, lo, L, tmp);
        for (int c = 0; c < 6; c++) permute_d(l->ct[c], so, 2 * L, tmp);
        if (linkOrder) memcpy(linkOrder, lo, sizeof(int64_t) * (size_t)L);
        free(lo);
        free(so);
    }
    if (newOrder) memcpy(newOrder, order, sizeof(int64_t) * (size_t)n);
    free(tmp);
    free(order);
    free(inv);
    return 0;
}

void hdmo_grid_free(hdmo_grid *g)
{
    free(g->colours); free(g->order); free(g->occColour); free(g->bucketStart); free(g->bucketSize);
    memset(g, 0, sizeof(*g));
}

/* OffloadHashMap::getValue replacement: bucket index or -1 (notFound). */
static int64_t bucket_of(const hdmo_grid *g, int64_t colour)
{
    int64_t lo = 0, hi = g->nOcc - 1;
    while (lo <= hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        int64_t c = g->occColour[mid];
        if (c == colour) return mid;
        if (c < colour) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* NeighAccel::iterateBlocks (NeighAccel.h:97-137) visiting
 * NeighbourGrid::iteratePossibleNeighbours (NeighbourGrid.h:85-105).
 * out == NULL: count only.  Returns the number of accepted neighbours. */
static int32_t neigh_visit(const hdmo_particles *p, const hdmo_grid *g, double interactDist,
                           int64_t pInd, int32_t *out)
{
    int32_t neighbours = 0;
    int colour = (int)g->colours[pInd];             /* passed as int (NeighbourGrid.h:87) */
    double xCord = p->x[pInd], yCord = p->y[pInd], zCord = p->z[pInd];
    for (int nn = 0; nn < 27; ++nn) {
        int targetColour = (int)(colour + g->dirOffsets[nn]);   /* NeighbourGrid.h:91 */
        int64_t bkt = bucket_of(g, targetColour);
        if (bkt < 0) continue;
        int32_t s = g->bucketStart[bkt], e = s + g->bucketSize[bkt];
        for (int32_t i = s; i < e; ++i) {
            int32_t nInd = g->order[i];
            if (nInd == pInd) continue;
            double xDist = xCord - p->x[nInd];
            double xDist2 = xDist * xDist;
            double yDist = yCord - p->y[nInd];
            double yDist2 = yDist * yDist;
            double zDist = zCord - p->z[nInd];
            double zDist2 = zDist * zDist;
            double particleDistance2 = xDist2 + yDist2 + zDist2;
            double particleDistance = sqrt(particleDistance2);
            double radSum = p->radius[pInd] + p->radius[nInd];
            double dist = particleDistance - radSum;
            if (dist < interactDist) {
                if (out) out[neighbours] = nInd;
                neighbours++;
            }
        }
    }
    return neighbours;
}

/* NeighAccel::generateNeighbours (NeighAccel.h:25-57): count pass, fill pass.
 * Stored as CSR; the reference's padded column-major ELL holds the same
 * per-particle sequences. */
int hdmo_neigh_build(const hdmo_particles *p, const hdmo_grid *g, double interactDist, hdmo_neigh *nl)
{
    int64_t n = p->n;
    memset(nl, 0, sizeof(*nl));
    nl->n = n;
    nl->count = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    nl->start = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    if (!nl->count || !nl->start) return -2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) nl->count[i] = neigh_visit(p, g, interactDist, i, NULL);
    nl->start[0] = 0;
    for (int64_t i = 0; i < n; i++) nl->start[i + 1] = nl->start[i] + nl->count[i];
    nl->list = (int32_t *)malloc(sizeof(int32_t) * (nl->start[n] ? nl->start[n] : 1));
    if (!nl->list) return -2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) neigh_visit(p, g, interactDist, i, nl->list + nl->start[i]);
    return 0;
}

void hdmo_neigh_free(hdmo_neigh *nl)
{
    free(nl->count); free(nl->start); free(nl->list);
    memset(nl, 0, sizeof(*nl));
}

/* ------------------------------------------------------------------------ */
/* Per-particle link lists (LinkAccelSide::generateLinks, :92-143, and       */
/* generatePerParticleIndexes, :43-67).  Appending links in ascending id     */
/* order gives the same ascending lists the reference's bubble sort          */
/* (:123-137) produces.                                                      */
/* ------------------------------------------------------------------------ */
int hdmo_link_lists(int64_t nParticles, const hdmo_links *l, hdmo_linklists *ll)
{
    memset(ll, 0, sizeof(*ll));
    ll->n = nParticles;
    ll->count = (int32_t *)calloc((size_t)(nParticles ? nParticles : 1), sizeof(int32_t));
    ll->start = (int64_t *)malloc(sizeof(int64_t) * (nParticles + 1));
    if (!ll->count || !ll->start) return -2;
    for (int64_t k = 0; k < l->n; k++) {
        if (l->a[k] < 0 || l->a[k] >= nParticles || l->b[k] < 0 || l->b[k] >= nParticles) return -3; /* dangling (LinkData.h:101-111) */
        ll->count[l->a[k]]++;
        ll->count[l->b[k]]++;
    }
    ll->start[0] = 0;
    for (int64_t i = 0; i < nParticles; i++) ll->start[i + 1] = ll->start[i] + ll->count[i];
    int64_t tot = ll->start[nParticles];
    ll->linkId = (int32_t *)malloc(sizeof(int32_t) * (tot ? tot : 1));
    ll->other = (int32_t *)malloc(sizeof(int32_t) * (tot ? tot : 1));
    ll->side = (int32_t *)malloc(sizeof(int32_t) * (tot ? tot : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nParticles ? nParticles : 1));
    if (!ll->linkId || !ll->other || !ll->side || !fill) return -2;
    for (int64_t i = 0; i < nParticles; i++) fill[i] = ll->start[i];
    for (int64_t k = 0; k < l->n; k++) {
        int32_t ends[2] = {l->a[k], l->b[k]};
        for (int e = 0; e < 2; e++) {
            int32_t pInd = ends[e];
            int64_t pos = fill[pInd]++;
            ll->linkId[pos] = (int32_t)k;
            /* LinkAccelSide.h:52-59 */
            int32_t other = l->b[k];
            int32_t side;
            if (other == pInd) {
                other = l->a[k];
                side = (int32_t)(2 * k + 1);
            } else {
                side = (int32_t)(2 * k);
            }
            ll->other[pos] = other;
            ll->side[pos] = side;
        }
    }
    free(fill);
    return 0;
}

void hdmo_linklists_free(hdmo_linklists *ll)
{
    free(ll->count); free(ll->start); free(ll->linkId); free(ll->other); free(ll->side);
    memset(ll, 0, sizeof(*ll));
}

/* ------------------------------------------------------------------------ */
/* Force functors                                                            */
/* ------------------------------------------------------------------------ */

/* HiDEMDefNeighForceFunc::operator() (HiDEMDefaultFunctions.h:38-86).
 * self/other = (xDiff, yDiff, zDiff, uDiff, vDiff, wDiff, radius). */
void hdmo_contact_pair(double mask, double xDist, double yDist, double zDist,
                       const double self[7], const double other[7],
                       const hdmo_force_params *fp, double force[6])
{
    double xDist2 = xDist * xDist;
    double yDist2 = yDist * yDist;
    double zDist2 = zDist * zDist;

    double particleDistance2 = xDist2 + yDist2 + zDist2;

    double radSum = self[6] + other[6];
    double activeDist2 = (radSum + fp->attractDistance) * (radSum + fp->attractDistance);

    mask = (particleDistance2 >= activeDist2) ? 0.0 : mask;

    double particleDistance = sqrt(particleDistance2);

    double invDist = 1.0 / particleDistance;
    double xComp = (xDist) * invDist;
    double yComp = (yDist) * invDist;
    double zComp = (zDist) * invDist;

    double Cnst = 0;
    double dimDist = (radSum - particleDistance);
    Cnst = (particleDistance < radSum) ? sqrt(dimDist * dimDist * dimDist) : Cnst;
    Cnst = (particleDistance >= radSum && particleDistance < radSum + fp->attractDistance)
               ? fp->attractScaleForce * dimDist : Cnst;
    Cnst *= fp->scaleE;

    force[0] += mask * ((-fp->damping) * (((self[0] - other[0]) * fp->invDt)) + (Cnst * xComp));
    force[1] += mask * ((-fp->damping) * (((self[1] - other[1]) * fp->invDt)) + (Cnst * yComp));
    force[2] += mask * ((-fp->damping) * (((self[2] - other[2]) * fp->invDt)) + (Cnst * zComp));
    force[3] += mask * ((-fp->dampingRotation) * (((self[3] - other[3]) * fp->invDt)));
    force[4] += mask * ((-fp->dampingRotation) * (((self[4] - other[4]) * fp->invDt)));
    force[5] += mask * ((-fp->dampingRotation) * (((self[5] - other[5]) * fp->invDt)));
}

/* HiDEMDefLinkForceFunc::operator() (HiDEMDefaultFunctions.h:105-203).
 * ct[] is read and updated in place (CT0..CT5 of this particle's side). */
void hdmo_beam_side(double mask, double xDist, double yDist, double zDist,
                    const double self[7], const double other[7],
                    const hdmo_force_params *fp, double ct[6], double force[6])
{
    const double invDt = fp->invDt, damping = fp->damping, dampingRotation = fp->dampingRotation;
    const double E = fp->E, W2 = fp->W2, W4 = fp->W4, G = fp->G;
    const double xDiff = self[0], yDiff = self[1], zDiff = self[2];
    const double uDiff = self[3], vDiff = self[4], wDiff = self[5];
    const double xDiffOther = other[0], yDiffOther = other[1], zDiffOther = other[2];
    const double uDiffOther = other[3], vDiffOther = other[4], wDiffOther = other[5];

    force[0] += (-damping * ((xDiff - xDiffOther) * invDt)) * mask;
    force[1] += (-damping * ((yDiff - yDiffOther) * invDt)) * mask;
    force[2] += (-damping * ((zDiff - zDiffOther) * invDt)) * mask;
    force[3] += (-dampingRotation * ((uDiff - uDiffOther) * invDt)) * mask;
    force[4] += (-dampingRotation * ((vDiff - vDiffOther) * invDt)) * mask;
    force[5] += (-dampingRotation * ((wDiff - wDiffOther) * invDt)) * mask;

    double xDist2 = xDist * xDist;
    double yDist2 = yDist * yDist;
    if (xDist == 0 && yDist == 0) {
        xDist2 = 1e-10;
        yDist2 = 1e-10;
    }
    double zDist2 = zDist * zDist;
    double distSum = xDist2 + yDist2 + zDist2;
    double invLI = 1.0 / sqrt(distSum);
    double invLP = 1.0 / sqrt(xDist2 + yDist2);
    double invLILP = invLI * invLP;
    double LI = invLI * (distSum);

    double T00 = xDist * invLI;
    double T01 = yDist * invLI;
    double T02 = zDist * invLI;

    double T10 = -yDist * invLP;
    double T11 = xDist * invLP;

    double T20 = -(zDist * xDist) * invLILP;
    double T21 = -(zDist * yDist) * invLILP;
    double T22 = (xDist2 + yDist2) * invLILP;

    double B[12];
    B[0] = T00 * (xDiff) + T01 * (yDiff) + T02 * (zDiff);
    B[1] = T10 * (xDiff) + T11 * (yDiff);
    B[2] = T20 * (xDiff) + T21 * (yDiff) + T22 * (zDiff);
    B[4] = T10 * (uDiff) + T11 * (vDiff);
    B[5] = T20 * (uDiff) + T21 * (vDiff) + T22 * (wDiff);

    B[6] = T00 * (xDiffOther) + T01 * (yDiffOther) + T02 * (zDiffOther);
    B[7] = T10 * (xDiffOther) + T11 * (yDiffOther);
    B[8] = T20 * (xDiffOther) + T21 * (yDiffOther) + T22 * (zDiffOther);
    B[10] = T10 * (uDiffOther) + T11 * (vDiffOther);
    B[11] = T20 * (uDiffOther) + T21 * (vDiffOther) + T22 * (wDiffOther);

    double L2 = LI * LI;
    double L3 = LI * LI * LI;

    double K[6];
    double LIDiv6 = (1.0 / 6.0) * invLI;

    K[0] = E * W2 * invLI;
    K[1] = E * W4 / (L3);
    K[5] = E * W4 / (2.0 * L2);

    ct[0] += K[0] * B[0] + (-K[0]) * B[6];
    ct[1] += K[1] * B[1] + K[5] * B[5] + (-K[1]) * B[7] + K[5] * B[11];
    ct[2] += K[1] * B[2] + (-K[5]) * B[4] + (-K[1]) * B[8] + (-K[5]) * B[10];

    force[0] += (-(T00 * ct[0] + T10 * ct[1] + T20 * ct[2])) * mask;
    force[1] += (-(T01 * ct[0] + T11 * ct[1] + T21 * ct[2])) * mask;
    force[2] += (-(T02 * ct[0] + T22 * ct[2])) * mask;

    K[2] = G * W4 * LIDiv6;
    K[3] = E * W4 * LIDiv6 * 2.0;
    K[4] = E * W4 * LIDiv6;

    B[3] = T00 * (uDiff) + T01 * (vDiff) + T02 * (wDiff);
    B[9] = T00 * (uDiffOther) + T01 * (vDiffOther) + T02 * (wDiffOther);

    ct[3] += K[2] * B[3] + (-K[2]) * B[9];
    ct[4] += (-K[5]) * B[2] + K[3] * B[4] + K[5] * B[8] + K[4] * B[10];
    ct[5] += K[5] * B[1] + K[3] * B[5] + (-K[5]) * B[7] + K[4] * B[11];

    force[3] += (-(T00 * ct[3] + T10 * ct[4] + T20 * ct[5])) * mask;
    force[4] += (-(T01 * ct[3] + T11 * ct[4] + T21 * ct[5])) * mask;
    force[5] += (-(T02 * ct[3] + T22 * ct[5])) * mask;
}

/* CombinedAccel::iterateCombined (CombinedAccel.h:121-197): per particle,
 * zero the force, apply the contact functor to every neighbour-list entry in
 * list order (:144-164), then the beam functor to every own link side in
 * ascending link id (:166-189, mask = state != 0), write CT back (:187) and
 * overwrite the particle's force (:192). */
void hdmo_force(hdmo_particles *p, hdmo_links *l, const hdmo_neigh *nl,
                const hdmo_linklists *ll, const hdmo_force_params *fp)
{
    int64_t n = p->n;
#pragma omp parallel for schedule(static)
    for (int64_t pInd = 0; pInd < n; pInd++) {
        double xCord = p->x[pInd], yCord = p->y[pInd], zCord = p->z[pInd];
        double force[6] = {0, 0, 0, 0, 0, 0};
        double self[7] = {p->dx[pInd], p->dy[pInd], p->dz[pInd], p->du[pInd], p->dv[pInd], p->dw[pInd], p->radius[pInd]};
        for (int64_t k = nl->start[pInd]; k < nl->start[pInd] + nl->count[pInd]; k++) {
            int32_t nInd = nl->list[k];
            double other[7] = {p->dx[nInd], p->dy[nInd], p->dz[nInd], p->du[nInd], p->dv[nInd], p->dw[nInd], p->radius[nInd]};
            double mask = 1;
            double xDist = xCord - p->x[nInd];
            double yDist = yCord - p->y[nInd];
            double zDist = zCord - p->z[nInd];
            hdmo_contact_pair(mask, xDist, yDist, zDist, self, other, fp, force);
        }
        for (int64_t k = ll->start[pInd]; k < ll->start[pInd] + ll->count[pInd]; k++) {
            int32_t lInd = ll->linkId[k];
            int32_t ctInd = ll->side[k];
            int32_t pIndOther = ll->other[k];
            double other[7] = {p->dx[pIndOther], p->dy[pIndOther], p->dz[pIndOther], p->du[pIndOther],
                               p->dv[pIndOther], p->dw[pIndOther], p->radius[pIndOther]};
            double xDist = xCord - p->x[pIndOther];
            double yDist = yCord - p->y[pIndOther];
            double zDist = zCord - p->z[pIndOther];
            double mask = 1;
            if (l->state[lInd] == 0) mask = 0;
            double ct[6];
            for (int c = 0; c < 6; c++) ct[c] = l->ct[c][ctInd];
            hdmo_beam_side(mask, xDist, yDist, zDist, self, other, fp, ct, force);
            for (int c = 0; c < 6; c++) l->ct[c][ctInd] = ct[c];
        }
        p->fx[pInd] = force[0]; p->fy[pInd] = force[1]; p->fz[pInd] = force[2];
        p->fu[pInd] = force[3]; p->fv[pInd] = force[4]; p->fw[pInd] = force[5];
    }
}

/* ------------------------------------------------------------------------ */
/* Time integration                                                          */
/* ------------------------------------------------------------------------ */

/* Damped central-difference update shared by both samples
 * (HiDEMDefTimeIntegration, HiDEMDefaultFunctions.h:317-370, identical to the
 * inline IceBlock form MainIceBlock.cpp:353-385). Returns the velocities. */
/* ---- 2-D maps (Map2d.h) --------------------------------------------------- */
void hdmo_map_init(hdmo_map2d *m)
{
    /* Map2d.h:56-59, float arithmetic */
    m->xRes = (m->xMax - m->xMin) / (float)(m->xPoints - 1);
    m->yRes = (m->yMax - m->yMin) / (float)(m->yPoints - 1);
    m->invXMaxMin = 1.0f / (m->xMax - m->xMin);
    m->invYMaxMin = 1.0f / (m->yMax - m->yMin);
}

/* floor of a float as an int, for indexes that the clamp below bounds: values
 * beyond [-1, n] (and NaN) clamp to those, so the conversion stays defined. */
static int32_t map_floor(float f, int32_t n)
{
    f = floorf(f);
    if (!(f >= -1.0f)) f = -1.0f;
    if (f > (float)n) f = (float)n;
    return (int32_t)f;
}

static int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

void hdmo_map_indexes(const hdmo_map2d *m, float x, float y, int32_t out[4])
{
    /* Map2d.h:96-106 */
    int32_t xPoint = map_floor(((x - m->xMin) * m->invXMaxMin) * (float)(m->xPoints - 1), m->xPoints);
    int32_t yPoint = map_floor(((y - m->yMin) * m->invYMaxMin) * (float)(m->yPoints - 1), m->yPoints);
    out[0] = clampi(xPoint, 0, m->xPoints - 1);
    out[1] = clampi(yPoint, 0, m->yPoints - 1);
    out[2] = clampi(xPoint + 1, 0, m->xPoints - 1);
    out[3] = clampi(yPoint + 1, 0, m->yPoints - 1);
}

double hdmo_map_interp(const hdmo_map2d *m, float x, float y)
{
    /* Map2d.h:110-145 */
    x += m->xOffset;
    y += m->yOffset;
    int32_t ix[4];
    hdmo_map_indexes(m, x, y, ix);
    int32_t xLow = ix[0], yLow = ix[1], xHigh = ix[2], yHigh = ix[3];
    if (xLow == xHigh && yLow == yHigh) return m->data[xLow + yLow * m->xPoints];
    float v00 = (float)m->data[xLow + yLow * m->xPoints];
    float v01 = (float)m->data[xLow + yHigh * m->xPoints];
    float v10 = (float)m->data[xHigh + yLow * m->xPoints];
    float v11 = (float)m->data[xHigh + yHigh * m->xPoints];
    float x0 = x - m->xMin;
    float y0 = y - m->yMin;
    float x1 = (float)xLow * m->xRes;
    float x2 = (float)xHigh * m->xRes;
    float y1 = (float)yLow * m->yRes;
    float y2 = (float)yHigh * m->yRes;
    float invX21 = 1.0f / (x2 - x1);
    float c1 = (x2 - x0) * invX21;
    float c2 = (x0 - x1) * invX21;
    float fxy1 = (c1 * v00) + (c2 * v10);
    float fxy2 = (c1 * v01) + (c2 * v11);
    if (xLow == xHigh) {
        fxy1 = v00;
        fxy2 = v01;
    } else if (yLow == yHigh) {
        return fxy1;
    }
    float invY21 = 1.0f / (y2 - y1);
    return (((y2 - y0) * invY21) * fxy1) + (((y0 - y1) * invY21) * fxy2);
}

static void central_difference(double dragCoeff, double mass, double inertia, const hdmo_integ_params *ip,
                               const double F[6], const double diff[6], double cord[3], double cordPrev[3],
                               double theta[3], double thetaPrev[3], double vel[6])
{
    double dragInvDt2 = dragCoeff * ip->invDt2;
    double massDtSqr = mass * ip->invDtSqr;
    double pMassDtSqr = mass / ip->dtSqr;
    double pMemDtSqr = (inertia / ip->dtSqr);
    double invC1 = 1.0 / (pMassDtSqr + dragCoeff * ip->invDt2);
    double invC2 = 1.0 / (pMemDtSqr + dragCoeff * ip->invDt2);
    double interDtSqr = inertia * ip->invDtSqr;

    double next[6];
    for (int c = 0; c < 3; c++) next[c] = (F[c] + massDtSqr * diff[c] + dragInvDt2 * -diff[c]) * invC1;
    for (int c = 3; c < 6; c++) next[c] = (F[c] + interDtSqr * diff[c] + dragInvDt2 * -diff[c]) * invC2;

    for (int c = 0; c < 3; c++) { cordPrev[c] = cord[c]; thetaPrev[c] = theta[c]; }
    for (int c = 0; c < 3; c++) cord[c] = cord[c] + next[c];
    for (int c = 0; c < 3; c++) theta[c] = theta[c] + next[3 + c];
    for (int c = 0; c < 3; c++) vel[c] = (next[c]) * ip->invDt;
    for (int c = 0; c < 3; c++) vel[3 + c] = (theta[c] - thetaPrev[c]) * ip->invDt;
}

/* Integrates every particle.  variant 0 = IceBlock applyForce
 * (MainIceBlock.cpp:299-400) with the build-defined clamp (type bit 0: the
 * particle does not move and its Diff is zeroed); variant 1 = GlacierDem
 * applyForce (MainDem.cpp:145-259) with an analytic bed (build-defined stand-in
 * for the InputBed maps), constant external force, HiDEMDefVelocityLimit
 * (HiDEMDefaultFunctions.h:244-290) and HiDEMDefBoundaryCondition (:292-315). */
void hdmo_integrate(hdmo_particles *p, const hdmo_integ_params *ip)
{
    int64_t n = p->n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (ip->variant == 0 && p->type && (p->type[i] & HDMO_TYPE_FIXED)) {
            p->dx[i] = 0; p->dy[i] = 0; p->dz[i] = 0;
            p->du[i] = 0; p->dv[i] = 0; p->dw[i] = 0;
            continue;
        }
        double cord[3] = {p->x[i], p->y[i], p->z[i]};
        double theta[3] = {p->tx[i], p->ty[i], p->tz[i]};
        double diff[6] = {p->dx[i], p->dy[i], p->dz[i], p->du[i], p->dv[i], p->dw[i]};
        double F[6] = {p->fx[i], p->fy[i], p->fz[i], p->fu[i], p->fv[i], p->fw[i]};
        double cordPrev[3], thetaPrev[3], vel[6];
        double radius = p->radius[i];
        double diameter = radius * 2.0;
        double zCord = cord[2];
        double dragCoeff;
        double particleMass = ip->rhoIce * diameter * diameter * diameter;
        double particleMomentInertia = particleMass * diameter * diameter / 6.0;
        double waterLevel = ip->blockWater ? ip->blockWater[ip->blockId[i]] : ip->waterLevel;

        if (ip->variant == 0) {
            if (zCord > waterLevel) dragCoeff = ip->airDrag; else dragCoeff = ip->waterDrag;
            if (zCord > waterLevel) F[2] += -ip->g * particleMass;
            if (zCord <= waterLevel) F[2] += ip->lift * particleMass;
            central_difference(dragCoeff, particleMass, particleMomentInertia, ip, F, diff, cord, cordPrev, theta, thetaPrev, vel);
        } else {
            /* InputBed::getHeight / getFric (InputGeometry.h:57-67) take the
             * coordinates as lpReal_t and Map2d::interp as float */
            double bedzCord = ip->maps[HDMO_MAP_BED]
                                  ? hdmo_map_interp(ip->maps[HDMO_MAP_BED], (float)cord[0], (float)cord[1])
                                  : ((cord[0] < ip->bedXEnd) ? ip->bedZ : ip->deepZ);
            double fric = ip->maps[HDMO_MAP_FRICTION]
                              ? hdmo_map_interp(ip->maps[HDMO_MAP_FRICTION], (float)cord[0], (float)cord[1])
                              : ((cord[0] < ip->frictionXEnd) ? 1.0 : 0.0);
            if ((zCord + radius) - bedzCord < ip->bedInteract) {
                double a = ip->bedDrag * fric;
                dragCoeff = (ip->frictionCap < a) ? ip->frictionCap : a;   /* std::min(a, cap) */
            } else if (zCord > waterLevel) {
                dragCoeff = ip->airDrag;
            } else {
                dragCoeff = ip->waterDrag;
            }
            if (zCord > waterLevel) F[2] += -ip->g * particleMass;
            if (zCord <= waterLevel) F[2] += ip->lift * particleMass;
            if (zCord + radius < bedzCord) F[2] += ip->bedStiffness * (bedzCord - (zCord + radius));
            /* ExternalForces maps (MainDem.cpp:210-216) */
            for (int c = 0; c < 6; c++) {
                const hdmo_map2d *m = ip->maps[HDMO_MAP_EXT + c];
                F[c] += m ? hdmo_map_interp(m, (float)cord[0], (float)cord[1]) : ip->ext[c];
            }
            central_difference(dragCoeff, particleMass, particleMomentInertia, ip, F, diff, cord, cordPrev, theta, thetaPrev, vel);
            if (ip->velocityLimit) {
                for (int c = 0; c < 3; c++) {
                    if (vel[c] > ip->maxVel) cord[c] = ip->dt * ip->maxVel + cordPrev[c];
                }
                for (int c = 0; c < 3; c++) {
                    if (vel[3 + c] > ip->maxRot) theta[c] = ip->dt * ip->maxRot + thetaPrev[c];
                }
                for (int c = 0; c < 3; c++) {
                    if (vel[c] < -ip->maxVel) cord[c] = ip->dt * -ip->maxVel + cordPrev[c];
                }
                for (int c = 0; c < 3; c++) {
                    if (vel[3 + c] < -ip->maxRot) theta[c] = ip->dt * -ip->maxRot + thetaPrev[c];
                }
            }
            for (int c = 0; c < 3; c++) {
                if (cord[c] + radius < ip->lower[c] && ip->bound[2 * c]) cord[c] = cordPrev[c];
                if (cord[c] - radius > ip->upper[c] && ip->bound[2 * c + 1]) cord[c] = cordPrev[c];
            }
        }

        p->x[i] = cord[0]; p->y[i] = cord[1]; p->z[i] = cord[2];
        p->tx[i] = theta[0]; p->ty[i] = theta[1]; p->tz[i] = theta[2];
        p->dx[i] = cord[0] - cordPrev[0];
        p->dy[i] = cord[1] - cordPrev[1];
        p->dz[i] = cord[2] - cordPrev[2];
        p->du[i] = theta[0] - thetaPrev[0];
        p->dv[i] = theta[1] - thetaPrev[1];
        p->dw[i] = theta[2] - thetaPrev[2];
    }
}

/* ------------------------------------------------------------------------ */
/* Link break: CombinedAccel::iterateLinks (CombinedAccel.h:200-235) over all */
/* links applying HiDEMDefLinkBreakFunc (HiDEMDefaultFunctions.h:217-241),    */
/* end 1 = localPartIndA, end 2 = localPartIndB.  abs() on lpReal_t is the    */
/* floating-point absolute value.                                             */
/* ------------------------------------------------------------------------ */
int64_t hdmo_break(const hdmo_particles *p, hdmo_links *l, const hdmo_break_params *bp)
{
    int64_t broken = 0;
#pragma omp parallel for schedule(static) reduction(+ : broken)
    for (int64_t k = 0; k < l->n; k++) {
        if (l->state[k] == 0) continue;
        int32_t a = l->a[k], b = l->b[k];
        double xDist = p->x[a] - p->x[b];
        double yDist = p->y[a] - p->y[b];
        double zDist = p->z[a] - p->z[b];
        double stretchLength = sqrt((xDist * xDist) + (yDist * yDist) + (zDist * zDist));
        double bx = fabs(p->tx[a] - p->tx[b]);
        double by = fabs(p->ty[a] - p->ty[b]);
        double bz = fabs(p->tz[a] - p->tz[b]);
        double inner = (by < bz) ? bz : by;               /* std::max(by, bz) */
        double maxBend = (bx < inner) ? inner : bx;       /* std::max(bx, inner) */
        double initialLen = l->initialLen[k];
        double load = (stretchLength - initialLen) / initialLen + bp->bendBreak * maxBend;
        double maxLoad = l->maxLoad ? l->maxLoad[k] : bp->maxLoad;
        if (load >= maxLoad) {
            l->state[k] = 0;
            broken++;
        }
    }
    return broken;
}

/* ------------------------------------------------------------------------ */
/* Rebuild policy                                                            */
/* ------------------------------------------------------------------------ */

/* HiDEMLib::getMaxDistOwned + accumulateDist (HiDEMLib.h:424-474). */
double hdmo_skin_heuristic(const hdmo_particles *p, int stepSize)
{
    double xMax = 0.0, yMax = 0.0, zMax = 0.0;  /* transform_reduce init 0.0 (:435) */
    for (int64_t i = 0; i < p->n; i++) {
        if (p->dx[i] > xMax) xMax = p->dx[i];
        if (p->dy[i] > yMax) yMax = p->dy[i];
        if (p->dz[i] > zMax) zMax = p->dz[i];
    }
    return stepSize * sqrt((xMax * xMax) + (yMax * yMax) + (zMax * zMax));
}

/* Exact Verlet-list displacement test (build-defined; no reference
 * counterpart): the largest displacement since the last list build. */
double hdmo_skin_exact(const hdmo_particles *p, const double *x0, const double *y0, const double *z0)
{
    double m2 = 0.0;
#pragma omp parallel for schedule(static) reduction(max : m2)
    for (int64_t i = 0; i < p->n; i++) {
        double ddx = p->x[i] - x0[i], ddy = p->y[i] - y0[i], ddz = p->z[i] - z0[i];
        double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        if (d2 > m2) m2 = d2;
    }
    return sqrt(m2);
}

/* ------------------------------------------------------------------------ */
/* Step driver: gridService (HiDEMLib.h:533-628) -> computePartForce ->       */
/* applyForce -> linkUpdate when ts*dt > breakTime -> finalizeTs              */
/* (MainIceBlock.cpp:409-447).  Redistribution and IO are not modelled.       */
/* ------------------------------------------------------------------------ */
int hdmo_sim_init(hdmo_sim *s)
{
    s->ts = 0;
    s->accumulatedDist = 0;
    s->rebuilds = 0;
    s->broken = 0;
    s->built = 0;
    memset(&s->grid, 0, sizeof(s->grid));
    memset(&s->nl, 0, sizeof(s->nl));
    memset(&s->ll, 0, sizeof(s->ll));
    int64_t n = s->p->n;
    s->x0 = (double *)malloc(sizeof(double) * (n ? n : 1));
    s->y0 = (double *)malloc(sizeof(double) * (n ? n : 1));
    s->z0 = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (!s->x0 || !s->y0 || !s->z0) return -2;
    return hdmo_link_lists(n, s->l, &s->ll);
}

static int sim_needs_rebuild(hdmo_sim *s)
{
    int regen = (s->built == 0);
    if (s->policy == 2) {
        if (s->regenInterval > 0 && (s->ts % s->regenInterval) == 0) regen = 1;
    } else if (s->policy == 0) {
        if (s->ts % 10 == 0) {
            s->accumulatedDist += hdmo_skin_heuristic(s->p, 10);
            if (s->accumulatedDist >= s->interactDist / 2.0) regen = 1;
        }
    } else {
        if (s->built) {
            double d = hdmo_skin_exact(s->p, s->x0, s->y0, s->z0);
            if (2.0 * d >= s->interactDist - s->fp.attractDistance) regen = 1;
        }
    }
    return regen;
}

int64_t hdmo_sim_run(hdmo_sim *s, int64_t nSteps, uint8_t *rebuildLog, int64_t *brokenLog, int64_t brokenCap)
{
    int64_t nLog = 0;
    uint8_t *prev = (uint8_t *)malloc((size_t)(s->l->n ? s->l->n : 1));
    for (int64_t step = 0; step < nSteps; step++) {
        int regen = sim_needs_rebuild(s);
        if (rebuildLog) rebuildLog[step] = (uint8_t)regen;
        if (regen) {
            s->accumulatedDist = 0;
            hdmo_grid_free(&s->grid);
            hdmo_neigh_free(&s->nl);
            hdmo_grid_build(s->p, s->searchDistance, &s->grid);
            hdmo_neigh_build(s->p, &s->grid, s->interactDist, &s->nl);
            memcpy(s->x0, s->p->x, sizeof(double) * s->p->n);
            memcpy(s->y0, s->p->y, sizeof(double) * s->p->n);
            memcpy(s->z0, s->p->z, sizeof(double) * s->p->n);
            s->built = 1;
            s->rebuilds++;
        }
        hdmo_force(s->p, s->l, &s->nl, &s->ll, &s->fp);
        hdmo_integrate(s->p, &s->ip);
        if ((double)s->ts * s->dt > s->breakTime) {
            if (brokenLog) memcpy(prev, s->l->state, (size_t)s->l->n);
            int64_t nb = hdmo_break(s->p, s->l, &s->bp);
            s->broken += nb;
            if (brokenLog && nb) {
                for (int64_t k = 0; k < s->l->n; k++) {
                    if (prev[k] != 0 && s->l->state[k] == 0) {
                        if (nLog < brokenCap) {
                            brokenLog[2 * nLog] = s->ts;
                            brokenLog[2 * nLog + 1] = k;
                        }
                        nLog++;
                    }
                }
            }
        }
        s->ts++;
    }
    free(prev);
    return nLog;
}

void hdmo_sim_free(hdmo_sim *s)
{
    hdmo_grid_free(&s->grid);
    hdmo_neigh_free(&s->nl);
    hdmo_linklists_free(&s->ll);
    free(s->x0); free(s->y0); free(s->z0);
    s->x0 = s->y0 = s->z0 = NULL;
}
