/*
 * CPU oracle for the multiphase-field time step -- TEST INFRASTRUCTURE ONLY (see mpf_oracle.h).
 *
 * Every pass cites the reference loop it restates. The grid has no halo: the periodic halo copies
 * of BoundaryConditions::SetX/Y/Z (include/BoundaryConditions.h:694-712, :825, :952) are replaced
 * by wrapped neighbour indices, which is what those copies produce for N >= Bcells.
 *
 * Compile without -march and with -ffp-contract=off so that no FMA contraction happens (the
 * reference's x86-64 build has no FMA either).
 */
#include "mpf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

typedef struct {
    uint32_t idx;
    double val;
    double lap;
    double grad[3];
} orc_entry;

typedef struct {
    int32_t n;
    int32_t flag;
    orc_entry e[ORC_CAP];
} orc_node;

/* antisymmetric pair increment (NodeAB<double,double>, include/Containers/NodeAB.h:37-69) */
typedef struct {
    uint32_t a, b;
    double v1, v2;
} orc_pair;

#define ORC_PCAP (ORC_CAP * ORC_CAP)

/* DrivingForceEntry (include/Containers/NodeDF.h:37-50): one per phase-field pair of an
 * interface node, created by UserDrivingForce::SetDrivingForce in list order */
typedef struct {
    uint32_t a, b;
    double raw, tmp, average, weight;
} orc_force;

typedef struct {
    int di, dj, dk;
    double w;          /* Laplacian weight / dx^2 */
    double gx, gy, gz; /* gradient weights / dx */
} orc_stencil;

struct orc_ctx {
    orc_params p;
    int nx, ny, nz;       /* storage sizes (reduced axis -> 1) */
    int dn[3];            /* active axes */
    int active;
    long ncells;
    int nphases;
    double* sigma;
    double* mu;
    int32_t* emodel;
    double* eps;
    double* dgval;
    int32_t* dgmode;
    int ngrains;
    int32_t* gphase;
    double* grot;
    uint8_t* gsolid;
    /* GrainsProperties (include/GrainsProperties.h:51-78) */
    double* volume;
    double* maxvolume;
    double* refvolume;
    double* vr;
    uint8_t* exist;
    int* stage; /* 0 seed, 1 nucleus, 2 stable (Grain::Clear sets Stable) */
    orc_node* nodes;
    orc_node* tmp;
    int nls, ngs;
    orc_stencil ls[27];
    orc_stencil gs[27];
    int need_grad;
    int64_t warnings;
    int max_list;   /* longest neighbour-merged list seen (entries incl. zero-valued) */
    int max_final;  /* longest finalized list seen */
    int status;
    double max_energy, max_mobility; /* over phase pairs of the last IP.Set */
    /* DrivingForce::Force of the current step (CSR over cells) and the averaging settings */
    int avg_on, avg_range, avg_weights;
    double avg_threshold;
    orc_force* force;
    long* force_off;
    long force_cap;
    /* per-step switches of the driving force (orc_set_step_terms) */
    int dg_step, avg_step;
    /* InterfaceRegularization: curvature entries of the current step (CSR over cells, one per pair
     * of non-zero entries of a wide-interface node, in the increments' pair order) and settings */
    int reg_on, reg_active, reg_range;
    double reg_threshold;
    orc_force* curv;
    long* curv_off;
    long curv_cap;
    /* lists the last step started from (orc_keep_step_start) */
    int keep_ip, ip_valid;
    orc_node* ipnodes;
};

static long cell_index(const orc_ctx* c, long i, long j, long k)
{
    return (i * c->ny + j) * c->nz + k;
}

static long wrap(long i, long n)
{
    long r = i % n;
    return r < 0 ? r + n : r;
}

/* Stencil tables: weights depend only on how many offset components are non-zero.
 * Laplacian: 1D three-point, 2D nine-point (Hale), 3D 27-point (Spotz-Carey);
 * gradient: 1D central, 2D eight-point and 3D 26-point (Alfaraj, Wang, Luo 2014).
 * Same numbers as include/Containers/LaplacianStencil.h:44-119 and GradientStencil.h:44-92;
 * enumeration order x outer, z inner, zero weights skipped (LaplacianStencil.h:197-219,
 * GradientStencil.h:132-156). */
static double lap_table(int active, int nnz)
{
    if (active == 1) return nnz == 0 ? -2.0 : (nnz == 1 ? 1.0 : 0.0);
    if (active == 2) return nnz == 0 ? -10.0 / 3.0 : (nnz == 1 ? 2.0 / 3.0 : 1.0 / 6.0);
    switch (nnz) {
    case 0: return -64.0 / 15.0;
    case 1: return 7.0 / 15.0;
    case 2: return 1.0 / 10.0;
    default: return 1.0 / 30.0;
    }
}

static double grad_table(int active, int nnz)
{
    if (active == 1) return nnz == 1 ? 0.5 : 0.0;
    if (active == 2) return nnz == 0 ? 0.0 : (nnz == 1 ? 1.0 / 3.0 : 0.25 / 3.0);
    switch (nnz) {
    case 0: return 0.0;
    case 1: return 1.0 / 4.64;
    case 2: return 0.245 / 4.64;
    default: return 0.085 / 4.64;
    }
}

static void build_stencils(orc_ctx* c)
{
    double dx = c->p.dx;
    c->nls = 0;
    c->ngs = 0;
    for (int x = -c->dn[0]; x <= c->dn[0]; ++x)
    for (int y = -c->dn[1]; y <= c->dn[1]; ++y)
    for (int z = -c->dn[2]; z <= c->dn[2]; ++z) {
        int nnz = (x != 0) + (y != 0) + (z != 0);
        double lw = lap_table(c->active, nnz);
        if (lw != 0.0) {
            orc_stencil* s = &c->ls[c->nls++];
            s->di = x; s->dj = y; s->dk = z;
            s->w = lw / (dx * dx);
        }
        double gw = grad_table(c->active, nnz);
        if (gw != 0.0) {
            orc_stencil* s = &c->gs[c->ngs++];
            s->di = x; s->dj = y; s->dk = z;
            s->gx = x * gw / dx;
            s->gy = y * gw / dx;
            s->gz = z * gw / dx;
        }
    }
}

/* PhaseField::CalculateReferenceVolume (src/PhaseField.cpp:367-391) with radius = iWidth */
static double reference_volume(const orc_ctx* c)
{
    double r = c->p.iwidth, v = 0.0;
    if (c->active == 1) v = r;
    else if (c->active == 2) v = ORC_PI * r * r;
    else v = (4.0 / 3.0) * ORC_PI * r * r * r;
    return v * c->p.nucleus_volume_factor;
}

static double* dup_d(const double* src, size_t n, double fill)
{
    double* d = (double*)malloc(n * sizeof(double));
    if (!d) return NULL;
    for (size_t i = 0; i < n; ++i) d[i] = src ? src[i] : fill;
    return d;
}

int orc_create(const orc_params* p, int32_t ngrains, const int32_t* grain_phase,
               const double* grain_rot, const uint8_t* grain_solid, orc_ctx** out)
{
    if (!p || !out || ngrains < 0 || p->nphases < 1 || p->dx <= 0.0) return ORC_ERR_ARG;
    if (p->nx < 0 || p->ny < 0 || p->nz < 0 || p->nx + p->ny + p->nz == 0) return ORC_ERR_ARG;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    if (!c) return ORC_ERR_NOMEM;
    c->p = *p;
    c->dn[0] = p->nx != 0; c->dn[1] = p->ny != 0; c->dn[2] = p->nz != 0;
    c->nx = p->nx ? p->nx : 1;
    c->ny = p->ny ? p->ny : 1;
    c->nz = p->nz ? p->nz : 1;
    c->active = c->dn[0] + c->dn[1] + c->dn[2];
    c->ncells = (long)c->nx * c->ny * c->nz;
    c->nphases = p->nphases;
    size_t np2 = (size_t)p->nphases * p->nphases;
    c->sigma = dup_d(p->sigma, np2, 0.0);
    c->mu = dup_d(p->mu, np2, 0.0);
    c->eps = dup_d(p->epsilon, np2, 0.0);
    c->dgval = p->dg_value ? dup_d(p->dg_value, np2, 0.0) : NULL;
    c->dgmode = (int32_t*)calloc(np2, sizeof(int32_t));
    for (size_t i = 0; i < np2 && c->dgmode; ++i) c->dgmode[i] = (p->dg_mode && p->dg_value) ? p->dg_mode[i] : 0;
    /* DrivingForce::Initialize / ReadInput (src/DrivingForce.cpp:45-90, 126-152) */
    c->avg_on = p->dg_average && p->dg_value && p->dg_mode;
    c->avg_range = p->dg_range > 0 ? p->dg_range : (int)p->iwidth;
    c->avg_threshold = p->dg_threshold > 0.0 ? p->dg_threshold : (p->iwidth < 5.0 ? p->iwidth / 15.0 : 1.0 / 3.0);
    c->avg_weights = p->dg_weights;
    if (c->avg_on) {
        /* the periodic halo of Range cells equals a wrapped index only for axes >= Range */
        if (c->avg_weights < 0 || c->avg_weights > 2 ||
            (c->dn[0] && c->nx < c->avg_range) || (c->dn[1] && c->ny < c->avg_range) ||
            (c->dn[2] && c->nz < c->avg_range)) {
            free(c);
            return ORC_ERR_ARG;
        }
        c->force_off = (long*)calloc((size_t)c->ncells + 1, sizeof(long));
    }
    c->dg_step = 1;
    c->avg_step = 1;
    /* InterfaceRegularization::Initialize / ReadInput (src/InterfaceRegularization.cpp:46-118) */
    c->reg_on = p->regularization != 0;
    c->reg_active = c->reg_on && (p->regularization_factor != 1.0 || p->curvature_factor != 1.0);
    c->reg_range = p->reg_range > 0 ? p->reg_range : (int)((p->iwidth + 1) / 2);
    c->reg_threshold = p->reg_threshold > 0.0 ? p->reg_threshold : DBL_EPSILON;
    if (c->reg_on) {
        if ((c->dn[0] && c->nx < c->reg_range) || (c->dn[1] && c->ny < c->reg_range) ||
            (c->dn[2] && c->nz < c->reg_range)) {
            free(c->force_off);
            free(c);
            return ORC_ERR_ARG;
        }
        c->curv_off = (long*)calloc((size_t)c->ncells + 1, sizeof(long));
    }
    c->emodel = (int32_t*)calloc(np2, sizeof(int32_t));
    c->need_grad = 0;
    for (size_t i = 0; i < np2; ++i) {
        c->emodel[i] = p->energy_model ? p->energy_model[i] : ORC_EN_ISO;
        if (c->emodel[i] != ORC_EN_ISO) c->need_grad = 1;
    }
    c->ngrains = ngrains;
    c->gphase = (int32_t*)calloc(ngrains ? ngrains : 1, sizeof(int32_t));
    c->grot = (double*)calloc(9 * (size_t)(ngrains ? ngrains : 1), sizeof(double));
    c->gsolid = (uint8_t*)calloc(ngrains ? ngrains : 1, 1);
    c->volume = (double*)calloc(ngrains ? ngrains : 1, sizeof(double));
    c->maxvolume = (double*)calloc(ngrains ? ngrains : 1, sizeof(double));
    c->refvolume = (double*)calloc(ngrains ? ngrains : 1, sizeof(double));
    c->vr = (double*)calloc(ngrains ? ngrains : 1, sizeof(double));
    c->exist = (uint8_t*)calloc(ngrains ? ngrains : 1, 1);
    c->stage = (int*)calloc(ngrains ? ngrains : 1, sizeof(int));
    c->nodes = (orc_node*)calloc((size_t)c->ncells, sizeof(orc_node));
    c->tmp = (orc_node*)calloc((size_t)c->ncells, sizeof(orc_node));
    if (!c->sigma || !c->mu || !c->eps || !c->emodel || !c->gphase || !c->grot || !c->nodes || !c->tmp ||
        (c->avg_on && !c->force_off) || (c->reg_on && !c->curv_off)) {
        orc_destroy(c);
        return ORC_ERR_NOMEM;
    }
    double refv = reference_volume(c);
    for (int g = 0; g < ngrains; ++g) {
        c->gphase[g] = grain_phase ? grain_phase[g] : 0;
        if (c->gphase[g] < 0 || c->gphase[g] >= p->nphases) { orc_destroy(c); return ORC_ERR_ARG; }
        for (int q = 0; q < 9; ++q)
            c->grot[9 * g + q] = grain_rot ? grain_rot[9 * g + q] : ((q % 4) == 0 ? 1.0 : 0.0);
        c->gsolid[g] = grain_solid ? grain_solid[g] : 1;
        /* PhaseField::AddGrainInfo (src/PhaseField.cpp:2970-2979) after GrainsProperties::add_grain */
        c->refvolume[g] = refv;
        c->vr[g] = 0.0;
        c->exist[g] = 1;
        c->stage[g] = 2;
    }
    build_stencils(c);
    *out = c;
    return ORC_OK;
}

void orc_destroy(orc_ctx* c)
{
    if (!c) return;
    free(c->sigma); free(c->mu); free(c->eps); free(c->emodel); free(c->dgval); free(c->dgmode);
    free(c->gphase); free(c->grot); free(c->gsolid);
    free(c->volume); free(c->maxvolume); free(c->refvolume); free(c->vr);
    free(c->exist); free(c->stage);
    free(c->nodes); free(c->tmp);
    free(c->force); free(c->force_off); free(c->ipnodes);
    free(c->curv); free(c->curv_off);
    free(c);
}

void orc_set_threads(int32_t n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_load(orc_ctx* c, const int32_t* count, const int32_t* index, const double* value)
{
    long off = 0;
    c->ip_valid = 0;
    for (long i = 0; i < c->ncells; ++i) {
        orc_node* nd = &c->nodes[i];
        if (count[i] < 0 || count[i] > ORC_CAP) return ORC_ERR_CAPACITY;
        nd->n = count[i];
        /* NodePF::read (include/Containers/NodePF.h:751-763): flag 2 when more than one entry */
        nd->flag = count[i] > 1 ? 2 : 0;
        for (int s = 0; s < count[i]; ++s, ++off) {
            if (index[off] < 0 || index[off] >= c->ngrains) return ORC_ERR_ARG;
            nd->e[s].idx = (uint32_t)index[off];
            nd->e[s].val = value[off];
            nd->e[s].lap = 0.0;
            nd->e[s].grad[0] = nd->e[s].grad[1] = nd->e[s].grad[2] = 0.0;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Pair list helpers (NodeAB semantics, include/Containers/NodeAB.h:321-547, 597-743)          */

static double ab_get_asym1(const orc_pair* d, int nd, uint32_t n, uint32_t m)
{
    for (int i = 0; i < nd; ++i) {
        if (d[i].a == n && d[i].b == m) return d[i].v1;
        if (d[i].a == m && d[i].b == n) return d[i].v1 * (-1.0);
    }
    return 0.0;
}

static double ab_get_asym2(const orc_pair* d, int nd, uint32_t n, uint32_t m)
{
    for (int i = 0; i < nd; ++i) {
        if (d[i].a == n && d[i].b == m) return d[i].v2;
        if (d[i].a == m && d[i].b == n) return d[i].v2 * (-1.0);
    }
    return 0.0;
}

static int ab_set_sym_pair(orc_pair* d, int* nd, uint32_t n, uint32_t m, double v1, double v2)
{
    for (int i = 0; i < *nd; ++i)
        if ((d[i].a == n && d[i].b == m) || (d[i].a == m && d[i].b == n)) {
            d[i].v1 = v1;
            d[i].v2 = v2;
            return 0;
        }
    if (*nd >= ORC_PCAP) return -1;
    d[*nd].a = n; d[*nd].b = m; d[*nd].v1 = v1; d[*nd].v2 = v2;
    ++*nd;
    return 0;
}

static int ab_add_asym1(orc_pair* d, int* nd, uint32_t n, uint32_t m, double v)
{
    for (int i = 0; i < *nd; ++i) {
        if (d[i].a == n && d[i].b == m) { d[i].v1 += v; return 0; }
        if (d[i].a == m && d[i].b == n) { d[i].v1 += v * (-1.0); return 0; }
    }
    if (*nd >= ORC_PCAP) return -1;
    d[*nd].a = n; d[*nd].b = m; d[*nd].v1 = v; d[*nd].v2 = 0.0;
    ++*nd;
    return 0;
}

static double std_min(double a, double b) { return (b < a) ? b : a; }
static double std_max(double a, double b) { return (a < b) ? b : a; }

static double node_get_value(const orc_node* nd, uint32_t idx)
{
    for (int i = 0; i < nd->n; ++i) if (nd->e[i].idx == idx) return nd->e[i].val;
    return 0.0;
}

/* NodePF::add_value (include/Containers/NodePF.h:288-306) */
static int node_add_value(orc_node* nd, uint32_t idx, double v)
{
    for (int i = 0; i < nd->n; ++i)
        if (nd->e[i].idx == idx) { nd->e[i].val += v; return 0; }
    if (nd->n >= ORC_CAP) return -1;
    if (nd->n) nd->flag = 2;
    orc_entry* e = &nd->e[nd->n++];
    e->idx = idx; e->val = v; e->lap = 0.0; e->grad[0] = e->grad[1] = e->grad[2] = 0.0;
    return 0;
}

/* InterfaceOrientation (src/InterfaceProperties.cpp:39-45): R^T * normalize(-grad) */
static void interface_orientation(const orc_ctx* c, const orc_entry* e, double n[3])
{
    double v[3] = {e->grad[0] * (-1.0), e->grad[1] * (-1.0), e->grad[2] * (-1.0)};
    double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len > DBL_EPSILON) {
        double inv = 1.0 / len;
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
    const double* R = &c->grot[9 * e->idx];
    for (int i = 0; i < 3; ++i) n[i] = R[0 * 3 + i] * v[0] + R[1 * 3 + i] * v[1] + R[2 * 3 + i] * v[2];
}

/* InterfaceEnergyModel::Calculate, Cubic branch (include/InterfaceProperties/InterfaceEnergyModel.h:325-345) */
static double cubic_energy(double energy, double eps1, const double n[3])
{
    return energy * (1.0 + eps1 * (1.5 - 2.5 * (pow(n[0], 4) + pow(n[1], 4) + pow(n[2], 4))));
}

static double max_energy_of(const orc_ctx* c, int pa, int pb)
{
    int k = pa * c->nphases + pb;
    if (c->emodel[k] == ORC_EN_CUBIC) return c->sigma[k] * (1.0 + c->eps[k] * 1.41);
    return c->sigma[k];
}

/* Interface energy InterfaceProperties::SetSR gives the pair (alpha = ea, beta = eb) of a
 * wide-interface node (src/InterfaceProperties.cpp:620-661): the phase-pair energy (ISO), or the
 * cubic energy of each solid grain's interface normal weighted by phi (1 - phi), divided by the
 * sum of the weights, MaxEnergy when neither grain weighs (CUBIC). */
static double ip_pair_energy(const orc_ctx* c, const orc_entry* ea, const orc_entry* eb)
{
    int pa = c->gphase[ea->idx], pb = c->gphase[eb->idx];
    int k = pa * c->nphases + pb;
    if (c->emodel[k] == ORC_EN_ISO) return c->sigma[k];
    double en = 0.0, wa = 0.0, wb = 0.0;
    if (c->gsolid[ea->idx]) {
        wa = fabs(ea->val * (1.0 - ea->val));
        if (wa > DBL_EPSILON) {
            double nrm[3];
            interface_orientation(c, ea, nrm);
            en += cubic_energy(c->sigma[k], c->eps[k], nrm) * wa;
        } else {
            wa = 0.0;
        }
    }
    if (c->gsolid[eb->idx]) {
        wb = fabs(eb->val * (1.0 - eb->val));
        if (wb > DBL_EPSILON) {
            double nrm[3];
            interface_orientation(c, eb, nrm);
            en += cubic_energy(c->sigma[k], c->eps[k], nrm) * wb;
        } else {
            wb = 0.0;
        }
    }
    if (wa + wb > DBL_EPSILON) en /= wa + wb;
    else en = max_energy_of(c, pa, pb);
    return en;
}

/* PhaseField::NormalizeIncrementsSR (src/PhaseField.cpp:1351-1567) on one node: the single-pair
 * branch (:1368-1404), zeroing of pairs that push a field at 0 below or at 1 above (:1407-1431),
 * removal of zero pairs, the iterative limiter with its 25-pass cap (:1451-1519), removal of
 * increments below DBL_EPSILON (:1521-1533) and the plausibility check, whose warnings are counted
 * (:1536-1561). */
static int normalize_pairs(const orc_node* nd, orc_pair* dot, int* nd_dot, double dt, int64_t* warnings)
{
    const int n = nd->n;
    if ((*nd_dot) == 1) {
        uint32_t ia = dot[0].a;
        double va = (dot[0].v1 + dot[0].v2) * dt;
        double old = node_get_value(nd, ia);
        if ((old == 0.0 && va < 0.0) || (old == 1.0 && va > 0.0)) {
            (*nd_dot) = 0;
        } else {
            double nv = old + va;
            double norm = 1.0;
            if (nv < 0.0) norm *= -old / va;
            else if (nv > 1.0) norm *= (1.0 - old) / va;
            if (norm > DBL_EPSILON) {
                dot[0].v1 *= norm;
                dot[0].v2 *= norm;
            } else {
                (*nd_dot) = 0;
            }
        }
    } else if ((*nd_dot) > 1) {
        for (int a = 0; a < n; ++a) {
            double dpsi = 0.0;
            for (int b = 0; b < n; ++b) {
                if (a == b) continue;
                dpsi += ab_get_asym1(dot, (*nd_dot), nd->e[a].idx, nd->e[b].idx);
                dpsi += ab_get_asym2(dot, (*nd_dot), nd->e[a].idx, nd->e[b].idx);
            }
            if ((nd->e[a].val == 0.0 && dpsi < 0.0) || (nd->e[a].val == 1.0 && dpsi > 0.0)) {
                for (int b = 0; b < n; ++b) {
                    if (a == b) continue;
                    if (ab_set_sym_pair(dot, &(*nd_dot), nd->e[a].idx, nd->e[b].idx, 0.0, 0.0)) {
                        return ORC_ERR_CAPACITY;
                    }
                }
            }
        }
        int w = 0;
        for (int i = 0; i < (*nd_dot); ++i)
            if (!(dot[i].v1 == 0.0 && dot[i].v2 == 0.0)) dot[w++] = dot[i];
        (*nd_dot) = w;

        if ((*nd_dot)) {
            /* per-index positive/negative sums and limits keyed by phase-field index */
            double pos[ORC_CAP], neg[ORC_CAP], lim1[ORC_CAP], lim2[ORC_CAP];
            int iter = 0;
            int needed = 1;
            while (needed) {
                ++iter;
                needed = 0;
                for (int a = 0; a < n; ++a) { pos[a] = 0.0; neg[a] = 0.0; }
                for (int i = 0; i < (*nd_dot); ++i) {
                    int sa = -1, sb = -1;
                    for (int a = 0; a < n; ++a) {
                        if (nd->e[a].idx == dot[i].a) sa = a;
                        if (nd->e[a].idx == dot[i].b) sb = a;
                    }
                    if (dot[i].v1 < 0.0) {
                        neg[sa] += dot[i].v1;
                        pos[sb] += -dot[i].v1;
                    }
                    if (dot[i].v1 > 0.0) {
                        pos[sa] += dot[i].v1;
                        neg[sb] += -dot[i].v1;
                    }
                }
                for (int a = 0; a < n; ++a) {
                    double phi = nd->e[a].val;
                    double npf = phi + (pos[a] + neg[a]) * dt;
                    lim2[a] = 1.0;
                    lim1[a] = 1.0;
                    if (npf < 0.0) lim2[a] = std_min(lim2[a], -(phi + pos[a] * dt) / (neg[a] * dt));
                    if (npf > 1.0) lim1[a] = std_min(lim1[a], (1.0 - (phi + neg[a] * dt)) / (pos[a] * dt));
                }
                for (int i = 0; i < (*nd_dot); ++i) {
                    int sa = -1, sb = -1;
                    for (int a = 0; a < n; ++a) {
                        if (nd->e[a].idx == dot[i].a) sa = a;
                        if (nd->e[a].idx == dot[i].b) sb = a;
                    }
                    if (dot[i].v1 < 0.0) {
                        double t = std_min(lim2[sa], lim1[sb]);
                        dot[i].v1 *= t;
                        if (t < 1.0) needed = 1;
                    }
                    if (dot[i].v1 > 0.0) {
                        double t = std_min(lim1[sa], lim2[sb]);
                        dot[i].v1 *= t;
                        if (t < 1.0) needed = 1;
                    }
                }
                if (iter > 24) needed = 0;
            }
            w = 0;
            for (int i = 0; i < (*nd_dot); ++i)
                if (!(fabs(dot[i].v1 * dt) < DBL_EPSILON)) dot[w++] = dot[i];
            (*nd_dot) = w;

            /* plausibility check: counts the warnings the reference prints */
            double tv[ORC_CAP];
            for (int a = 0; a < n; ++a) tv[a] = nd->e[a].val;
            for (int i = 0; i < (*nd_dot); ++i) {
                if (dot[i].v1 == 0.0) continue;
                for (int a = 0; a < n; ++a) {
                    if (nd->e[a].idx == dot[i].a) tv[a] += dot[i].v1 * dt;
                    if (nd->e[a].idx == dot[i].b) tv[a] += -dot[i].v1 * dt;
                }
            }
            for (int a = 0; a < n; ++a)
                if (tv[a] < -FLT_EPSILON || tv[a] > 1.0 + FLT_EPSILON) ++*warnings;
        }
    }

    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* One node: IP.Set -> DO increments (-> driving force) -> NormalizeIncrements -> MergeIncrements */

typedef struct {
    double max_energy, max_mobility;
    int64_t warnings;
    int err;
} local_result;

/* loc_increment of pair (a, b) of a node (src/DoubleObstacle.cpp:517-562, the same terms at
 * :702-752): the pair term, then per third entry the energy-difference term and the
 * triple-junction term, before the mobility */
static double loc_increment(const orc_ctx* c, const orc_node* nd, double sig[ORC_CAP][ORC_CAP], int a, int b,
                            double scale, double prefactor)
{
    const int n = nd->n;
    const orc_entry* ea = &nd->e[a];
    const orc_entry* eb = &nd->e[b];
    double d = sig[a][b] * ((ea->lap + 0.5 * (1.0 + scale) * prefactor * ea->val) -
                            (eb->lap + 0.5 * (1.0 + scale) * prefactor * eb->val));
    if (n > 2)
    for (int g = 0; g < n; ++g) {
        if (g == a || g == b) continue;
        const orc_entry* eg = &nd->e[g];
        d += (sig[b][g] - sig[a][g]) * (eg->lap + 0.5 * (1.0 + scale) * prefactor * eg->val);
        if (c->p.tj_model == ORC_TJ_MODEL) {
            if (c->p.tj_factor != 0.0 || scale < 1.0) {
                int pa = c->gphase[ea->idx], pb = c->gphase[eb->idx], pg = c->gphase[eg->idx];
                int np = c->nphases;
                d += std_max(c->p.tj_factor, (1.0 - scale)) * prefactor *
                     (c->sigma[pa * np + pb] + c->sigma[pb * np + pg] + c->sigma[pa * np + pg]) *
                     (ea->val * eg->val - eb->val * eg->val);
            }
        } else if (c->p.tj_model == ORC_TJ_VALUE) {
            d += c->p.tj_energy * prefactor * (ea->val * eg->val - eb->val * eg->val);
        }
    }
    return d;
}

/* InterfaceProperties::SetSR energies and mobilities of a node's pairs
 * (src/InterfaceProperties.cpp:603-762): ISO and CUBIC energy, ISO mobility; a pair with both 0
 * stays 0 (Properties(i,j,k) is cleared first). */
static void pair_tables(const orc_ctx* c, const orc_node* nd, double sig[ORC_CAP][ORC_CAP],
                        double mob[ORC_CAP][ORC_CAP], double* max_energy, double* max_mobility)
{
    const int n = nd->n;
    memset(sig, 0, sizeof(double) * ORC_CAP * ORC_CAP);
    memset(mob, 0, sizeof(double) * ORC_CAP * ORC_CAP);
    for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b) {
        const orc_entry* ea = &nd->e[a];
        const orc_entry* eb = &nd->e[b];
        int pa = c->gphase[ea->idx], pb = c->gphase[eb->idx];
        int k = pa * c->nphases + pb;
        double en = ip_pair_energy(c, ea, eb);
        double mo = c->mu[k];
        if (en != 0.0 || mo != 0.0) {
            sig[a][b] = sig[b][a] = en;
            mob[a][b] = mob[b][a] = mo;
        }
        *max_energy = std_max(*max_energy, en);
        *max_mobility = std_max(*max_mobility, mo);
    }
}

static void local_step(orc_ctx* c, orc_node* nd, double dt, local_result* res)
{
    const int n = nd->n;
    double sig[ORC_CAP][ORC_CAP];
    double mob[ORC_CAP][ORC_CAP];
    pair_tables(c, nd, sig, mob, &res->max_energy, &res->max_mobility);

    /* DoubleObstacle::CalculatePhaseFieldIncrementsSR (src/DoubleObstacle.cpp:502-571), or with
     * InterfaceRegularization the curvature overload (:686-770): the same loc_increment, scaled left
     * to right (loc_increment * mobility * norm_1 * scale) and then by RegularizationFactor */
    orc_pair dot[ORC_PCAP];
    int nd_dot = 0;
    const double eta = c->p.iwidth * c->p.dx;
    const double prefactor = ORC_PI * ORC_PI / (eta * eta);
    double localn = 0.0;
    if (c->p.consider_nucleus_volume) {
        for (int a = 0; a < n; ++a) localn += c->vr[nd->e[a].idx];
    } else {
        localn = (double)n;
    }
    const double norm_1 = 1.0 / localn;
    for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b) {
        const orc_entry* ea = &nd->e[a];
        const orc_entry* eb = &nd->e[b];
        double scale = sqrt(c->vr[ea->idx] * c->vr[eb->idx]);
        double d = loc_increment(c, nd, sig, a, b, scale, prefactor);
        if (c->reg_on) {
            d = d * mob[a][b] * norm_1 * scale;
            if (c->reg_active) d *= c->p.regularization_factor;
        } else {
            d *= mob[a][b] * norm_1 * scale;
        }
        if (ab_add_asym1(dot, &nd_dot, ea->idx, eb->idx, d)) { res->err = ORC_ERR_CAPACITY; return; }
    }

    /* InterfaceRegularization::MergePhaseFieldIncrementsSR (src/InterfaceRegularization.cpp:425-456)
     * on interface nodes: (average or raw) * mobility * weight * Pi/Eta per curvature entry. The
     * weight is 0 unless Average (SetWeightsSR) ran this step. */
    if (c->reg_active && nd->flag == 2) {
        const double pref = ORC_PI / eta;
        const long ci = nd - c->nodes;
        for (long q = c->curv_off[ci]; q < c->curv_off[ci + 1]; ++q) {
            const orc_force* f = &c->curv[q];
            int ia = -1, ib = -1;
            for (int t = 0; t < n; ++t) {
                if (nd->e[t].idx == f->a) ia = t;
                if (nd->e[t].idx == f->b) ib = t;
            }
            const double kappa = c->p.reg_average ? f->average : f->raw;
            const double dpsi = kappa * mob[ia][ib] * f->weight * pref;
            if (ab_add_asym1(dot, &nd_dot, f->a, f->b, dpsi)) { res->err = ORC_ERR_CAPACITY; return; }
        }
    }

    /* UserDrivingForce::SetDrivingForce (src/UserDrivingForce.cpp:152-175) on interface nodes,
     * DrivingForce::SkipAverage (src/DrivingForce.cpp:257-274) or the averaged value of the
     * node's Force entry (dg_average_pass), and DrivingForce::MergePhaseFieldIncrementsSR
     * (src/DrivingForce.cpp:497-602), Limiting on. No grain is a Seed (grains enter Stable and
     * only nucleation makes seeds), so the 1e-6 seed term of :528-536 never applies. */
    if (c->dgval && c->dg_step && nd->flag == 2) {
        const double pref = ORC_PI / eta;
        const orc_force* fe = c->avg_on && c->avg_step ? &c->force[c->force_off[nd - c->nodes]] : NULL;
        for (int a = 0; a < n - 1; ++a)
        for (int b = a + 1; b < n; ++b) {
            int pa = c->gphase[nd->e[a].idx], pb = c->gphase[nd->e[b].idx];
            if (c->dgmode[pa * c->nphases + pb] != 1) continue; /* no Force entry for this pair */
            double raw = c->dgval[pa * c->nphases + pb];
            if (fe) raw = (fe++)->average;
            double weight = sqrt(nd->e[a].val * nd->e[b].val);
            double allowed = c->p.dg_limit * pref * max_energy_of(c, pa, pb);
            allowed *= c->p.regularization_factor;
            double dg = allowed * tanh(raw / allowed);
            double dpsi = dg * mob[a][b] * weight * pref;
            if (ab_add_asym1(dot, &nd_dot, nd->e[a].idx, nd->e[b].idx, dpsi)) { res->err = ORC_ERR_CAPACITY; return; }
        }
    }

    /* PhaseField::NormalizeIncrementsSR (src/PhaseField.cpp:1351-1567) */
    {
        int st = normalize_pairs(nd, dot, &nd_dot, dt, &res->warnings);
        if (st) { res->err = st; return; }
    }

    /* PhaseField::MergeIncrementsSR (src/PhaseField.cpp:1268-1299) */
    for (int i = 0; i < nd_dot; ++i) {
        double value = 1.0 * (dot[i].v1 + dot[i].v2) * dt;
        if (fabs(value) >= DBL_EPSILON) {
            if (node_add_value(nd, dot[i].a, value) || node_add_value(nd, dot[i].b, -value)) {
                res->err = ORC_ERR_CAPACITY;
                return;
            }
        }
    }
}

/* NodePF::finalize (include/Containers/NodePF.h:214-255) */
static int node_finalize(orc_node* nd)
{
    double total = 0.0;
    int w = 0;
    for (int i = 0; i < nd->n; ++i) {
        orc_entry e = nd->e[i];
        int erase = 0;
        if (e.val >= 1.0 - DBL_EPSILON) e.val = 1.0;
        if (e.val <= 0.0 + DBL_EPSILON) {
            e.val = 0.0;
            erase = 1;
        }
        total += e.val;
        if (!erase) nd->e[w++] = e;
    }
    nd->n = w;
    if (w == 0) return -1;
    if (w > 1) {
        double norm = 1.0 / total;
        for (int i = 0; i < w; ++i) {
            nd->e[i].val *= norm;
            nd->e[i].lap = 0.0;
            nd->e[i].grad[0] = nd->e[i].grad[1] = nd->e[i].grad[2] = 0.0;
        }
        nd->flag = 2;
    } else {
        nd->e[0].val = 1.0;
        nd->e[0].lap = 0.0;
        nd->e[0].grad[0] = nd->e[0].grad[1] = nd->e[0].grad[2] = 0.0;
        nd->flag = 0;
    }
    return 0;
}

/* PhaseField::FinalizeSR (src/PhaseField.cpp:1008-1033) */
static int finalize_sr(orc_ctx* c)
{
    int err = 0;
    const long N = c->ncells;
#pragma omp parallel for schedule(static) reduction(|:err)
    for (long i = 0; i < N; ++i)
        if (c->nodes[i].flag != 0)
            if (node_finalize(&c->nodes[i])) err |= 1;
    if (err) return ORC_ERR_EMPTY_NODE;

    /* SetFlagsSR (src/PhaseField.cpp:953-969): flag 1 on the 3^d neighbours of flag-2 nodes */
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ++ci) {
        long i = ci / ((long)c->ny * c->nz), j = (ci / c->nz) % c->ny, k = ci % c->nz;
        int f = c->nodes[ci].flag;
        int nf = f;
        if (f == 0) {
            for (int ii = -c->dn[0]; ii <= c->dn[0] && !nf; ++ii)
            for (int jj = -c->dn[1]; jj <= c->dn[1] && !nf; ++jj)
            for (int kk = -c->dn[2]; kk <= c->dn[2] && !nf; ++kk) {
                long nb = cell_index(c, wrap(i + ii, c->nx), wrap(j + jj, c->ny), wrap(k + kk, c->nz));
                if (c->nodes[nb].flag == 2) nf = 1;
            }
        }
        c->tmp[ci].flag = nf;
    }
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ++ci) c->nodes[ci].flag = c->tmp[ci].flag;

    /* CalculateDerivativesSR (src/PhaseField.cpp:566-615): Laplacian and gradient of every
     * neighbour entry with a non-zero value, accumulated into a copy of the own list; indices
     * not yet present are appended with value 0 (the neighbour-list merge). */
#pragma omp parallel for schedule(static) reduction(|:err)
    for (long ci = 0; ci < N; ++ci) {
        orc_node* dst = &c->tmp[ci];
        const orc_node* src = &c->nodes[ci];
        *dst = *src;
        if (src->flag == 0) continue;
        long i = ci / ((long)c->ny * c->nz), j = (ci / c->nz) % c->ny, k = ci % c->nz;
        for (int s = 0; s < c->nls; ++s) {
            const orc_stencil* st = &c->ls[s];
            long nb = cell_index(c, wrap(i + st->di, c->nx), wrap(j + st->dj, c->ny), wrap(k + st->dk, c->nz));
            const orc_node* nn = &c->nodes[nb];
            for (int q = 0; q < nn->n; ++q) {
                if (nn->e[q].val == 0.0) continue;
                double lap = st->w * nn->e[q].val;
                int found = 0;
                for (int t = 0; t < dst->n; ++t)
                    if (dst->e[t].idx == nn->e[q].idx) { dst->e[t].lap += lap; found = 1; break; }
                if (!found) {
                    if (dst->n >= ORC_CAP) { err |= 1; break; }
                    orc_entry* e = &dst->e[dst->n++];
                    e->idx = nn->e[q].idx; e->val = 0.0; e->lap = lap;
                    e->grad[0] = e->grad[1] = e->grad[2] = 0.0;
                }
            }
        }
        if (c->need_grad)
        for (int s = 0; s < c->ngs; ++s) {
            const orc_stencil* st = &c->gs[s];
            long nb = cell_index(c, wrap(i + st->di, c->nx), wrap(j + st->dj, c->ny), wrap(k + st->dk, c->nz));
            const orc_node* nn = &c->nodes[nb];
            for (int q = 0; q < nn->n; ++q) {
                if (nn->e[q].val == 0.0) continue;
                double gx = st->gx * nn->e[q].val, gy = st->gy * nn->e[q].val, gz = st->gz * nn->e[q].val;
                int found = 0;
                for (int t = 0; t < dst->n; ++t)
                    if (dst->e[t].idx == nn->e[q].idx) {
                        dst->e[t].grad[0] = dst->e[t].grad[0] + gx;
                        dst->e[t].grad[1] = dst->e[t].grad[1] + gy;
                        dst->e[t].grad[2] = dst->e[t].grad[2] + gz;
                        found = 1;
                        break;
                    }
                if (!found) {
                    if (dst->n >= ORC_CAP) { err |= 1; break; }
                    orc_entry* e = &dst->e[dst->n++];
                    e->idx = nn->e[q].idx; e->val = 0.0; e->lap = 0.0;
                    e->grad[0] = gx; e->grad[1] = gy; e->grad[2] = gz;
                }
            }
        }
    }
    if (err) return ORC_ERR_CAPACITY;
    orc_node* t = c->nodes; c->nodes = c->tmp; c->tmp = t;
    {
        int mx = 0, mf = 0;
        for (long ci = 0; ci < N; ++ci) {
            const orc_node* nd = &c->nodes[ci];
            if (nd->n > mx) mx = nd->n;
            int nz = 0;
            for (int q = 0; q < nd->n; ++q) nz += nd->e[q].val != 0.0;
            if (nz > mf) mf = nz;
        }
        c->max_list = mx > c->max_list ? mx : c->max_list;
        c->max_final = mf > c->max_final ? mf : c->max_final;
    }

    /* CalculateGrainsVolume (src/PhaseField.cpp:812-951), single-thread summation order */
    for (int g = 0; g < c->ngrains; ++g) c->volume[g] = 0.0;
    for (long ci = 0; ci < N; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        for (int q = 0; q < nd->n; ++q) c->volume[nd->e[q].idx] += nd->e[q].val;
    }
    for (int g = 0; g < c->ngrains; ++g) {
        c->maxvolume[g] = std_max(c->volume[g], c->maxvolume[g]);
        c->vr[g] = c->maxvolume[g] / c->refvolume[g];
    }
    for (int g = 0; g < c->ngrains; ++g) {
        if (c->exist[g] && c->volume[g] <= 0.0) {
            c->exist[g] = 0;
            c->stage[g] = 2;
            c->volume[g] = 0.0;
            c->maxvolume[g] = 0.0;
            c->vr[g] = 1.0;
        } else if (c->volume[g] > 0.0) {
            c->exist[g] = 1;
            if (c->stage[g] == 0) c->stage[g] = 1;
            if (c->vr[g] > 1.0) {
                c->stage[g] = 2;
                c->vr[g] = 1.0;
            }
        }
    }
    return ORC_OK;
}

int orc_finalize_initialization(orc_ctx* c)
{
    c->ip_valid = 0;
    return finalize_sr(c);
}

/* ------------------------------------------------------------------------------------------ */
/* Driving-force averaging over the interface                                                  */

/* NodeDF::get_weight / get_raw / get_tmp (include/Containers/NodeDF.h:359-410): symmetric weight,
 * antisymmetric raw and tmp (sign flips with the pair order), 0 for a missing pair */
static const orc_force* force_find(const orc_ctx* c, long cell, uint32_t a, uint32_t b, double* sign)
{
    for (long q = c->force_off[cell]; q < c->force_off[cell + 1]; ++q) {
        const orc_force* f = &c->force[q];
        if (f->a == a && f->b == b) { *sign = 1.0; return f; }
        if (f->a == b && f->b == a) { *sign = -1.0; return f; }
    }
    return NULL;
}

/* UserDrivingForce::SetDrivingForce (src/UserDrivingForce.cpp:152-175) after DrivingForce::Clear
 * (src/DrivingForce.cpp:176-183): one entry per VALUE-mode pair of every interface node (the
 * whole list, zero-valued entries included), raw = Value(phaseA, phaseB); then
 * DrivingForce::SetWeights (:276-291): weight = sqrt(phiA*phiB). */
static int dg_set_forces(orc_ctx* c)
{
    const long N = c->ncells;
    const int np = c->nphases;
    c->force_off[0] = 0;
    for (long ci = 0; ci < N; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        long m = 0;
        if (nd->flag == 2)
            for (int a = 0; a < nd->n - 1; ++a)
            for (int b = a + 1; b < nd->n; ++b)
                m += c->dgmode[c->gphase[nd->e[a].idx] * np + c->gphase[nd->e[b].idx]] == 1;
        c->force_off[ci + 1] = c->force_off[ci] + m;
    }
    long total = c->force_off[N];
    if (total > c->force_cap) {
        orc_force* f = (orc_force*)realloc(c->force, (size_t)(total ? total : 1) * sizeof(orc_force));
        if (!f) return ORC_ERR_NOMEM;
        c->force = f;
        c->force_cap = total;
    }
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        orc_force* f = &c->force[c->force_off[ci]];
        if (nd->flag != 2) continue;
        for (int a = 0; a < nd->n - 1; ++a)
        for (int b = a + 1; b < nd->n; ++b) {
            int pa = c->gphase[nd->e[a].idx], pb = c->gphase[nd->e[b].idx];
            if (c->dgmode[pa * np + pb] != 1) continue;
            f->a = nd->e[a].idx; f->b = nd->e[b].idx;
            f->raw = c->dgval[pa * np + pb];
            f->tmp = 0.0; f->average = 0.0;
            f->weight = sqrt(nd->e[a].val * nd->e[b].val);
            ++f;
        }
    }
    return ORC_OK;
}

/* DrivingForce::CollectAverage (second == 0, src/DrivingForce.cpp:293-379) and
 * DistributeAverage (second == 1, :381-478): weighted mean over the sphere of radius Range
 * (offsets x outer, z inner, each clipped to N-1 on its axis) of the neighbours' raw (tmp)
 * values, over neighbours whose weight exceeds 0 (threshold*VR_A*VR_B). The periodic
 * BoundaryConditions copies of Force between the passes are wrapped indices here. */
static void dg_average_pass(orc_ctx* c, int second)
{
    const int R = c->avg_range;
    const int xr = (R < c->nx - 1 ? R : c->nx - 1) * c->dn[0];
    const int yr = (R < c->ny - 1 ? R : c->ny - 1) * c->dn[1];
    const int zr = (R < c->nz - 1 ? R : c->nz - 1) * c->dn[2];
    const double thr = sqrt(c->avg_threshold * (1.0 - c->avg_threshold));
    const long N = c->ncells;
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ++ci) {
        if (c->nodes[ci].flag != 2) continue;
        long i = ci / ((long)c->ny * c->nz), j = (ci / c->nz) % c->ny, k = ci % c->nz;
        for (long q = c->force_off[ci]; q < c->force_off[ci + 1]; ++q) {
            orc_force* e = &c->force[q];
            const double scale = c->vr[e->a] * c->vr[e->b];
            if (second ? !(e->weight > 0.0) : !(e->weight > thr * scale)) {
                if (second && (c->stage[e->a] == 0 || c->stage[e->b] == 0)) e->average = e->raw;
                continue;
            }
            double value = 0.0, sum = 0.0;
            int counter = 0;
            for (int ii = -xr; ii <= xr; ii++)
            for (int jj = -yr; jj <= yr; jj++)
            for (int kk = -zr; kk <= zr; kk++) {
                double wdist = R - sqrt((double)(ii * ii + jj * jj + kk * kk));
                long nb = cell_index(c, wrap(i + ii, c->nx), wrap(j + jj, c->ny), wrap(k + kk, c->nz));
                double sign = 1.0;
                const orc_force* f = force_find(c, nb, e->a, e->b, &sign);
                double wphi = f ? f->weight : 0.0;
                if (wdist > 0.0 && (second ? wphi > thr * scale : wphi > 0.0)) {
                    double w = 0.0;
                    switch (c->avg_weights) {
                    case 1: w = wdist; sum += w; break;
                    case 0: w = wphi; sum += w; break;
                    default: counter++; w = 1.0; break;
                    }
                    value += w * ((second ? f->tmp : f->raw) * sign);
                }
            }
            double fallback = second ? e->tmp : e->raw;
            double r;
            if (c->avg_weights == 2) r = counter != 0 ? value / counter : fallback;
            else r = sum > 0.0 ? value / sum : fallback;
            if (second) e->average = r;
            else e->tmp = r;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Interface regularization (curvature averaged over the interface)                             */

/* The curvature part of DoubleObstacle::CalculatePhaseFieldIncrementsSR with InterfaceRegularization
 * (src/DoubleObstacle.cpp:686-770) after InterfaceRegularization::Clear: on every wide-interface node,
 * per pair of its list with both values non-zero (pair order a < b),
 *   raw = loc_increment * Eta/Pi * scale * norm_1 / sqrt(phi_a phi_b) * (CurvatureFactor - RegularizationFactor)
 * (NodeDF::add_raw on a fresh entry); then SetWeightsSR (src/InterfaceRegularization.cpp:178-194) on
 * interface nodes, weight = sqrt(phi_a phi_b), when Average is on. */
static int curv_set_raw(orc_ctx* c)
{
    const long N = c->ncells;
    c->curv_off[0] = 0;
    for (long ci = 0; ci < N; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        long m = 0;
        if (nd->flag != 0)
            for (int a = 0; a < nd->n; ++a)
            for (int b = a + 1; b < nd->n; ++b)
                m += nd->e[a].val != 0.0 && nd->e[b].val != 0.0;
        c->curv_off[ci + 1] = c->curv_off[ci] + m;
    }
    long total = c->curv_off[N];
    if (total > c->curv_cap) {
        orc_force* f = (orc_force*)realloc(c->curv, (size_t)(total ? total : 1) * sizeof(orc_force));
        if (!f) return ORC_ERR_NOMEM;
        c->curv = f;
        c->curv_cap = total;
    }
    const double eta = c->p.iwidth * c->p.dx;
    const double prefactor = ORC_PI * ORC_PI / (eta * eta);
    const double prefactor2 = eta / ORC_PI;
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        if (nd->flag == 0) continue;
        orc_force* f = &c->curv[c->curv_off[ci]];
        double sig[ORC_CAP][ORC_CAP], mob[ORC_CAP][ORC_CAP], me = 0.0, mm = 0.0;
        pair_tables(c, nd, sig, mob, &me, &mm);
        double localn = 0.0;
        if (c->p.consider_nucleus_volume) {
            for (int a = 0; a < nd->n; ++a) localn += c->vr[nd->e[a].idx];
        } else {
            localn = (double)nd->n;
        }
        const double norm_1 = 1.0 / localn;
        for (int a = 0; a < nd->n; ++a)
        for (int b = a + 1; b < nd->n; ++b) {
            const orc_entry* ea = &nd->e[a];
            const orc_entry* eb = &nd->e[b];
            if (!(ea->val != 0.0 && eb->val != 0.0)) continue;
            const double scale = sqrt(c->vr[ea->idx] * c->vr[eb->idx]);
            const double li = loc_increment(c, nd, sig, a, b, scale, prefactor);
            double kappa = li * prefactor2 * scale * norm_1 / sqrt(ea->val * eb->val);
            kappa *= c->p.curvature_factor - c->p.regularization_factor;
            f->a = ea->idx; f->b = eb->idx;
            f->raw = kappa;
            f->tmp = 0.0; f->average = 0.0;
            f->weight = (c->p.reg_average && nd->flag == 2) ? sqrt(ea->val * eb->val) : 0.0;
            ++f;
        }
    }
    return ORC_OK;
}

static const orc_force* curv_find(const orc_ctx* c, long cell, uint32_t a, uint32_t b, double* sign)
{
    for (long q = c->curv_off[cell]; q < c->curv_off[cell + 1]; ++q) {
        const orc_force* f = &c->curv[q];
        if (f->a == a && f->b == b) { *sign = 1.0; return f; }
        if (f->a == b && f->b == a) { *sign = -1.0; return f; }
    }
    return NULL;
}

/* InterfaceRegularization::CollectAverageSR (second == 0, src/InterfaceRegularization.cpp:196-223)
 * and DistributeAverageSR (second == 1, :225-274) on interface nodes: over the sphere of radius
 * Range (offsets x outer, z inner, each clipped to N-1 on its axis; interface neighbours only),
 * tmp = sum(weight*(Range-d)*raw)/sum(weight*(Range-d)) over terms above DBL_EPSILON, for entries
 * whose weight exceeds sqrt(T(1-T)); average = the mean of the neighbours' tmp whose weight exceeds
 * sqrt(T(1-T)), for entries whose weight exceeds DBL_EPSILON. Seed grains keep raw. The periodic
 * BoundaryConditions copies of Curvature between the passes are wrapped indices here. */
static void curv_average_pass(orc_ctx* c, int second)
{
    const int R = c->reg_range;
    const int xr = (R < c->nx - 1 ? R : c->nx - 1) * c->dn[0];
    const int yr = (R < c->ny - 1 ? R : c->ny - 1) * c->dn[1];
    const int zr = (R < c->nz - 1 ? R : c->nz - 1) * c->dn[2];
    const double thr = sqrt(c->reg_threshold * (1.0 - c->reg_threshold));
    const long N = c->ncells;
#pragma omp parallel for schedule(static)
    for (long ci = 0; ci < N; ++ci) {
        if (c->nodes[ci].flag != 2) continue;
        long i = ci / ((long)c->ny * c->nz), j = (ci / c->nz) % c->ny, k = ci % c->nz;
        for (long q = c->curv_off[ci]; q < c->curv_off[ci + 1]; ++q) {
            orc_force* e = &c->curv[q];
            if (second ? !(e->weight > DBL_EPSILON) : !(e->weight > thr)) {
                if (c->stage[e->a] == 0 || c->stage[e->b] == 0) {
                    if (second) e->average = e->raw;
                    else e->tmp = e->raw;
                }
                continue;
            }
            double value = 0.0, sum = 0.0;
            int counter = 0;
            for (int ii = -xr; ii <= xr; ii++)
            for (int jj = -yr; jj <= yr; jj++)
            for (int kk = -zr; kk <= zr; kk++) {
                double wdist = R - sqrt((double)(ii * ii + jj * jj + kk * kk));
                long nb = cell_index(c, wrap(i + ii, c->nx), wrap(j + jj, c->ny), wrap(k + kk, c->nz));
                if (!(wdist > 0.0) || c->nodes[nb].flag != 2) continue;
                double sign = 1.0;
                const orc_force* f = curv_find(c, nb, e->a, e->b, &sign);
                const double wphi = f ? f->weight : 0.0;
                if (second) {
                    if (wphi > thr) {
                        counter++;
                        value += f->tmp * sign;
                    }
                } else {
                    const double w = wphi * wdist;
                    if (w > DBL_EPSILON) {
                        sum += w;
                        value += w * (f->raw * sign);
                    }
                }
            }
            if (second) e->average = counter != 0 ? value / counter : e->tmp;
            else if (sum > DBL_EPSILON) e->tmp = value / sum;
        }
    }
}

int orc_step(orc_ctx* c, double dt, int32_t nsteps)
{
    for (int s = 0; s < nsteps; ++s) {
        if (c->keep_ip) {
            /* the lists this step's IP.Set sees (neighbour-merged, with their derivatives) */
            memcpy(c->ipnodes, c->nodes, (size_t)c->ncells * sizeof(orc_node));
            c->ip_valid = 1;
        }
        if (c->avg_on && c->avg_step && c->dg_step) {
            /* DrivingForce::Average (src/DrivingForce.cpp:242-255) on the step-start lists */
            int st = dg_set_forces(c);
            if (st) return st;
            dg_average_pass(c, 0);
            dg_average_pass(c, 1);
        }
        if (c->reg_active) {
            /* DO.CalculatePhaseFieldIncrements(Phase, IP, Kappa)'s curvature, Kappa.Average
             * (src/InterfaceRegularization.cpp:156-176), on the step-start lists */
            int st = curv_set_raw(c);
            if (st) return st;
            if (c->p.reg_average) {
                curv_average_pass(c, 0);
                curv_average_pass(c, 1);
            }
        }
        double mx_e = 0.0, mx_m = 0.0;
        int64_t warn = 0;
        int err = 0;
        const long N = c->ncells;
#pragma omp parallel
        {
            local_result r = {0.0, 0.0, 0, 0};
#pragma omp for schedule(static)
            for (long i = 0; i < N; ++i)
                if (c->nodes[i].flag != 0) local_step(c, &c->nodes[i], dt, &r);
#pragma omp critical
            {
                mx_e = std_max(mx_e, r.max_energy);
                mx_m = std_max(mx_m, r.max_mobility);
                warn += r.warnings;
                err |= r.err;
            }
        }
        c->max_energy = mx_e;
        c->max_mobility = mx_m;
        c->warnings += warn;
        if (err) return err;
        int st = finalize_sr(c);
        if (st) return st;
    }
    return ORC_OK;
}

int64_t orc_total_entries(const orc_ctx* c)
{
    int64_t t = 0;
    for (long i = 0; i < c->ncells; ++i)
        for (int q = 0; q < c->nodes[i].n; ++q) t += c->nodes[i].e[q].val != 0.0;
    return t;
}

int orc_download(const orc_ctx* c, int32_t* count, int32_t* index, double* value, int8_t* flag)
{
    long off = 0;
    for (long i = 0; i < c->ncells; ++i) {
        const orc_node* nd = &c->nodes[i];
        int m = 0;
        for (int q = 0; q < nd->n; ++q) {
            if (nd->e[q].val == 0.0) continue;
            if (index) index[off] = (int32_t)nd->e[q].idx;
            if (value) value[off] = nd->e[q].val;
            ++off;
            ++m;
        }
        if (count) count[i] = m;
        if (flag) flag[i] = (int8_t)nd->flag;
    }
    return ORC_OK;
}

int orc_grains(const orc_ctx* c, double* volume, double* volume_ratio)
{
    for (int g = 0; g < c->ngrains; ++g) {
        if (volume) volume[g] = c->volume[g];
        if (volume_ratio) volume_ratio[g] = c->vr[g];
    }
    return ORC_OK;
}

int orc_grain_records(const orc_ctx* c, double* maxvolume, double* refvolume, uint8_t* exist, int32_t* stage)
{
    for (int g = 0; g < c->ngrains; ++g) {
        if (maxvolume) maxvolume[g] = c->maxvolume[g];
        if (refvolume) refvolume[g] = c->refvolume[g];
        if (exist) exist[g] = c->exist[g];
        if (stage) stage[g] = c->stage[g];
    }
    return ORC_OK;
}

int64_t orc_limiter_warnings(const orc_ctx* c) { return c->warnings; }

double orc_max_time_step(const orc_ctx* c)
{
    double dt = DBL_MAX;
    if (c->max_energy > DBL_EPSILON && c->max_mobility > 0.0) {
        double f = c->active == 1 ? 0.50 : (c->active == 2 ? 0.25 : 0.16);
        dt = f * c->p.dx * c->p.dx / (c->max_mobility * c->max_energy * c->p.regularization_factor);
    }
    return dt;
}

void orc_list_stats(const orc_ctx* c, int32_t* max_merged, int32_t* max_finalized)
{
    if (max_merged) *max_merged = c->max_list;
    if (max_finalized) *max_finalized = c->max_final;
}

int orc_normalize_pairs(int32_t n, const int32_t* idx, const double* val, int32_t* npairs_io, int32_t* pa, int32_t* pb,
                        double* v1, double dt, int64_t* warnings)
{
    const int npairs = *npairs_io;
    if (n < 0 || n > ORC_CAP || npairs < 0 || npairs > ORC_PCAP) return ORC_ERR_CAPACITY;
    orc_node nd;
    memset(&nd, 0, sizeof(nd));
    nd.n = n;
    nd.flag = 2;
    for (int i = 0; i < n; ++i) {
        nd.e[i].idx = (uint32_t)idx[i];
        nd.e[i].val = val[i];
    }
    orc_pair dot[ORC_PCAP];
    int nd_dot = 0;
    for (int i = 0; i < npairs; ++i)
        if (ab_add_asym1(dot, &nd_dot, (uint32_t)pa[i], (uint32_t)pb[i], v1[i])) return ORC_ERR_CAPACITY;
    int64_t w = 0;
    int st = normalize_pairs(&nd, dot, &nd_dot, dt, &w);
    if (st) return st;
    for (int i = 0; i < nd_dot; ++i) {
        pa[i] = (int32_t)dot[i].a;
        pb[i] = (int32_t)dot[i].b;
        v1[i] = dot[i].v1;
    }
    *npairs_io = nd_dot;
    if (warnings) *warnings = w;
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Driving-force switches and the interface energy                                            */

int orc_set_step_terms(orc_ctx* c, int32_t driving_force, int32_t average)
{
    if (!c) return ORC_ERR_ARG;
    c->dg_step = driving_force != 0;
    c->avg_step = average != 0;
    return ORC_OK;
}

int orc_keep_step_start(orc_ctx* c, int32_t on)
{
    if (!c) return ORC_ERR_ARG;
    if (on && !c->ipnodes) {
        c->ipnodes = (orc_node*)calloc((size_t)c->ncells, sizeof(orc_node));
        if (!c->ipnodes) return ORC_ERR_NOMEM;
    }
    c->keep_ip = on != 0;
    c->ip_valid = 0;
    return ORC_OK;
}

/* DoubleObstacle::PointEnergy (src/DoubleObstacle.cpp:903-923) summed in storage order, as
 * DoubleObstacle::Energy / AverageEnergyDensity (:953-967, :989-1005) sum it on one thread: over
 * every pair alpha < beta of a wide-interface node's list (zero-valued appended entries included)
 * 4 E (phi_a phi_b - Eta^2/Pi^2 grad_a . grad_b) / Eta, with the gradients of the current lists
 * (PhaseField::CalculateDerivativesSR, src/PhaseField.cpp:588-603) and E the energy
 * InterfaceProperties::Properties holds for the pair: what the last IP.Set gave it on the lists
 * of `ipn` (NULL: no IP.Set yet, every energy 0; a pair absent there, or a node that was not a
 * wide interface, reads 0, NodeIP::get_energy include/Containers/NodeIP.h:379-388). */
static double energy_sum(const orc_ctx* c, const orc_node* ipn)
{
    const double eta = c->p.iwidth * c->p.dx;
    const double pref = eta * eta / (ORC_PI * ORC_PI);
    double sum = 0.0;
    for (long ci = 0; ci < c->ncells; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        double e = 0.0;
        if (nd->flag != 0 && ipn) {
            double g[ORC_CAP][3];
            memset(g, 0, sizeof(g));
            long i = ci / ((long)c->ny * c->nz), j = (ci / c->nz) % c->ny, k = ci % c->nz;
            for (int s = 0; s < c->ngs; ++s) {
                const orc_stencil* st = &c->gs[s];
                long nb = cell_index(c, wrap(i + st->di, c->nx), wrap(j + st->dj, c->ny), wrap(k + st->dk, c->nz));
                const orc_node* nn = &c->nodes[nb];
                for (int q = 0; q < nn->n; ++q) {
                    if (nn->e[q].val == 0.0) continue;
                    for (int t = 0; t < nd->n; ++t)
                        if (nd->e[t].idx == nn->e[q].idx) {
                            g[t][0] = g[t][0] + st->gx * nn->e[q].val;
                            g[t][1] = g[t][1] + st->gy * nn->e[q].val;
                            g[t][2] = g[t][2] + st->gz * nn->e[q].val;
                            break;
                        }
                }
            }
            const orc_node* ip = &ipn[ci];
            for (int a = 0; a < nd->n; ++a)
            for (int b = a + 1; b < nd->n; ++b) {
                double en = 0.0;
                if (ip->flag != 0) {
                    int pa = -1, pb = -1;
                    for (int t = 0; t < ip->n; ++t) {
                        if (ip->e[t].idx == nd->e[a].idx) pa = t;
                        if (ip->e[t].idx == nd->e[b].idx) pb = t;
                    }
                    if (pa >= 0 && pb >= 0)
                        en = pa < pb ? ip_pair_energy(c, &ip->e[pa], &ip->e[pb]) : ip_pair_energy(c, &ip->e[pb], &ip->e[pa]);
                }
                const double dot = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
                e += 4.0 * en * (nd->e[a].val * nd->e[b].val - pref * dot) / eta;
            }
        }
        sum += e;
    }
    return sum;
}

int orc_interface_energy(const orc_ctx* c, int32_t ip_state, double* sum)
{
    if (!c || !sum) return ORC_ERR_ARG;
    const orc_node* ipn = NULL;
    if (ip_state == ORC_IP_CURRENT) ipn = c->nodes;
    else if (ip_state == ORC_IP_PREVIOUS) {
        if (!c->keep_ip || !c->ip_valid) return ORC_ERR_ARG;
        ipn = c->ipnodes;
    } else if (ip_state != ORC_IP_NONE) {
        return ORC_ERR_ARG;
    }
    *sum = energy_sum(c, ipn);
    return ORC_OK;
}

/* PhaseField::CalculateFractions (src/PhaseField.cpp:411-423): per cell and thermodynamic phase
 * the sum of its grains' values in list order, out[phase * ncells + cell]. */
int orc_phase_fractions(const orc_ctx* c, double* out)
{
    if (!c || !out) return ORC_ERR_ARG;
    const long N = c->ncells;
    for (long i = 0; i < (long)c->nphases * N; ++i) out[i] = 0.0;
    for (long ci = 0; ci < N; ++ci) {
        const orc_node* nd = &c->nodes[ci];
        for (int q = 0; q < nd->n; ++q) out[(long)c->gphase[nd->e[q].idx] * N + ci] += nd->e[q].val;
    }
    return ORC_OK;
}
