/*
 * xr_oracle.c -- CPU restatement of the XR density build and Hamiltonian
 * contractions.  TEST INFRASTRUCTURE ONLY: this file is the checker that the
 * HIP library is compared against.  Nothing in the product path may link or
 * call it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.
 *
 * Each function names the reference routine whose algorithm it follows.  The
 * loop nest of the density build (ket configuration -> operator indices,
 * rightmost operator outermost -> Pauli/core checks -> lexical rank -> rank-1
 * scatter over all state pairs) is kept, because the excitation-list order and
 * the sign convention are part of the contract:
 *   general-XRCC/density_tensors.c:29-62   (lexical rank of a valence config)
 *   general-XRCC/density_tensors.c:74-134  (delete/insert with parity, core check)
 *   general-XRCC/density_tensors.c:142-154 (rank-1 scatter into all (I,J))
 *   general-XRCC/density_tensors.c:162-556 (a, c, ca, aa, cc, caa, cca, ccaa)
 *   Be-states/_attic_/deprecated/density_tensors.c:560-688 (caaa, ccca, and the
 *       separate bra coefficient list of :142-154)
 *   general-XRCC/H_contractions.c:22-252   (scalar contractions)
 *
 * Sign convention: removing or inserting spin orbital p in an ascending list
 * of occupied spin orbitals gives (-1)^(number of occupied orbitals > p);
 * a Pauli violation gives 0.  Spin orbital p < n_orbs is alpha, p >= n_orbs
 * is beta; the first n_core of each spin are core and must stay occupied.
 *
 * Build: oracle/Makefile -> oracle/liboracle.so (gcc -O2, no GPU).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* ------------------------------------------------------------------------ */
/* binomial coefficients and the lexical rank                               */
/* ------------------------------------------------------------------------ */

static i64 binom(i64 n, i64 k)
{
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    i64 r = 1;
    for (i64 i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

/* Table of C(j, i) in row i, column j (zero below the diagonal), the layout of
 * FCIcomboMat (general-XRCC/build_density_tensors.py:31-37). */
void orc_combo_table(i64 n_e, i64 n_so, i64 *out)
{
    for (i64 i = 0; i < n_e; i++)
        for (i64 j = 0; j < n_so; j++)
            out[i * n_so + j] = (j >= i) ? binom(j, i) : 0;
}

/* Position of the ascending combination vc[0..ne) of {0..nso-1} in the
 * lexical enumeration (itertools.combinations order).  Counts, position by
 * position, the combinations that agree on the earlier entries and carry a
 * smaller value here; follows find_valence_config_index
 * (general-XRCC/density_tensors.c:29-47), reading C(m, r) from the FCIcomboMat
 * table (row r, column m) like the reference does. */
static i64 lex_rank(const i64 *vc, i64 ne, i64 nso, const i64 *comb)
{
    i64 rank = 0, prev = -1;
    for (i64 i = 0; i < ne; i++) {
        i64 rows_left = ne - 1 - i;
        for (i64 v = prev + 1; v < vc[i]; v++) {
            i64 m = nso - 1 - v;
            rank += comb[rows_left * nso + m];
        }
        prev = vc[i];
    }
    return rank;
}

/* Drops the core orbitals and shifts beta indices down, then ranks; follows
 * find_config_index (general-XRCC/density_tensors.c:49-62). */
i64 orc_config_rank(const i64 *cfg, i64 n_elec, i64 n_orbs, i64 n_core,
                    const i64 *comb)
{
    i64 nve = n_elec - 2 * n_core, nvso = 2 * (n_orbs - n_core);
    i64 val[128];
    i64 k = 0;
    for (i64 i = 0; i < n_elec; i++) {
        i64 p = cfg[i];
        if (p >= n_core && p < n_orbs) val[k++] = p - n_core;
        else if (p >= n_orbs + n_core && p < 2 * n_orbs) val[k++] = p - 2 * n_core;
    }
    if (nve <= 0) return 0;
    return lex_rank(val, nve, nvso, comb);
}

/* ------------------------------------------------------------------------ */
/* operators on ascending occupation lists                                  */
/* ------------------------------------------------------------------------ */

/* Removes p (follows annihilate, density_tensors.c:80-95).  Returns the
 * parity, or 0 if p is empty; list shrinks by one on success. */
static int remove_orb(i64 p, i64 *cfg, i64 n)
{
    i64 pos = -1;
    for (i64 i = 0; i < n; i++) if (cfg[i] == p) { pos = i; break; }
    if (pos < 0) return 0;
    i64 above = n - 1 - pos;
    for (i64 i = pos; i < n - 1; i++) cfg[i] = cfg[i + 1];
    return (above & 1) ? -1 : 1;
}

/* Inserts p (follows create, density_tensors.c:97-113).  cfg needs room for
 * n+1 entries.  Returns the parity, or 0 if p is already occupied. */
static int insert_orb(i64 p, i64 *cfg, i64 n)
{
    i64 pos = 0;
    while (pos < n && cfg[pos] < p) pos++;
    if (pos < n && cfg[pos] == p) return 0;
    for (i64 i = n; i > pos; i--) cfg[i] = cfg[i - 1];
    cfg[pos] = p;
    i64 above = n - pos;
    return (above & 1) ? -1 : 1;
}

/* All 2*n_core core spin orbitals present (all_core_occupied,
 * density_tensors.c:115-134). */
static int core_full(const i64 *cfg, i64 n, i64 n_orbs, i64 n_core)
{
    for (i64 c = 0; c < n_core; c++) {
        int ha = 0, hb = 0;
        for (i64 i = 0; i < n; i++) {
            if (cfg[i] == c) ha = 1;
            if (cfg[i] == n_orbs + c) hb = 1;
        }
        if (!ha || !hb) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* density build: one walker for every operator string                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    const char *ops;     /* e.g. "ccaa": ops[0] is the leftmost operator      */
    i64 k;               /* number of operators                                */
    i64 dim, n_orbs, n_core;
    i64 n_elec_bra;
    const i64 *comb_bra; /* FCIcomboMat of the bra space                       */
    /* density accumulation (mode 0) */
    double *out;
    i64 tsize;
    i64 n_bra_states, n_bra_cfg, n_ket_states, n_ket_cfg;
    const double *zbra, *zket;
    /* excitation-list emission (mode 1) or counting (mode 2) */
    int mode;
    i64 n_emit, cap;
    i64 *e_q, *e_idx, *e_p;
    int8_t *e_sign;
    i64 Q;
} walker;

static void scatter(walker *w, i64 P, i64 idx, int par)
{
    /* z_contract, density_tensors.c:142-154: every (I,J) pair gets the
     * rank-1 update  par * zbra[I,P] * zket[J,Q]. */
    for (i64 I = 0; I < w->n_bra_states; I++) {
        double zi = w->zbra[I * w->n_bra_cfg + P] * (double)par;
        double *row = w->out + (I * w->n_ket_states) * w->tsize + idx;
        for (i64 J = 0; J < w->n_ket_states; J++)
            row[J * w->tsize] += zi * w->zket[J * w->n_ket_cfg + w->Q];
    }
}

static void walk(walker *w, int level, const i64 *cfg, i64 n, int par, i64 *ind)
{
    if (level < 0) {
        if (!core_full(cfg, n, w->n_orbs, w->n_core)) return;
        i64 P = orc_config_rank(cfg, w->n_elec_bra, w->n_orbs, w->n_core, w->comb_bra);
        i64 idx = 0;
        for (i64 l = 0; l < w->k; l++) idx = idx * w->dim + ind[l];
        if (w->mode == 0) scatter(w, P, idx, par);
        else {
            if (w->mode == 1 && w->n_emit < w->cap) {
                w->e_q[w->n_emit] = w->Q;
                w->e_idx[w->n_emit] = idx;
                w->e_p[w->n_emit] = P;
                w->e_sign[w->n_emit] = (int8_t)par;
            }
            w->n_emit++;
        }
        return;
    }
    i64 next[130];
    int is_create = (w->ops[level] == 'c');
    for (i64 x = 0; x < w->dim; x++) {
        memcpy(next, cfg, (size_t)n * sizeof(i64));
        int s = is_create ? insert_orb(x, next, n) : remove_orb(x, next, n);
        if (s == 0) continue;
        ind[level] = x;
        walk(w, level - 1, next, is_create ? n + 1 : n - 1, par * s, ind);
    }
}

static void walker_init(walker *w, const char *ops, i64 n_orbs, i64 n_core,
                        i64 n_elec_ket, const i64 *comb_bra)
{
    memset(w, 0, sizeof(*w));
    w->ops = ops;
    w->k = (i64)strlen(ops);
    w->dim = 2 * n_orbs;
    w->n_orbs = n_orbs;
    w->n_core = n_core;
    i64 nb = n_elec_ket;
    for (i64 l = 0; l < w->k; l++) nb += (ops[l] == 'c') ? 1 : -1;
    w->n_elec_bra = nb;
    w->comb_bra = comb_bra;
}

/* Dense transition density  out[(I*nJ+J)*dim^k + idx] = <I| ops |J>  for the
 * operator string `ops` (characters 'c' and 'a', leftmost first).  out must be
 * zeroed by the caller (the reference accumulates, build_density_tensors.py:94).
 * zbra: (n_bra_states x n_bra_cfg), zket: (n_ket_states x n_ket_cfg), both
 * row-major; ket_cfgs: (n_ket_cfg x n_elec_ket) ascending occupation lists in
 * lexical order; comb_bra: FCIcomboMat of the bra valence space.  Follows the
 * *_tensor kernels (general-XRCC/density_tensors.c:162-556 and
 * Be-states/_attic_/deprecated/density_tensors.c:560-688). */
void orc_density(const char *ops, double *out,
                 i64 n_orbs, i64 n_core,
                 i64 n_bra_states, i64 n_bra_cfg, const double *zbra,
                 const i64 *comb_bra,
                 i64 n_ket_states, i64 n_ket_cfg, const double *zket,
                 i64 n_elec_ket, const i64 *ket_cfgs)
{
    walker w;
    walker_init(&w, ops, n_orbs, n_core, n_elec_ket, comb_bra);
    w.out = out;
    w.tsize = 1;
    for (i64 l = 0; l < w.k; l++) w.tsize *= w.dim;
    w.n_bra_states = n_bra_states;
    w.n_bra_cfg = n_bra_cfg;
    w.n_ket_states = n_ket_states;
    w.n_ket_cfg = n_ket_cfg;
    w.zbra = zbra;
    w.zket = zket;
    w.mode = 0;
    i64 ind[8];
    for (i64 Q = 0; Q < n_ket_cfg; Q++) {
        w.Q = Q;
        walk(&w, (int)w.k - 1, ket_cfgs + Q * n_elec_ket, n_elec_ket, 1, ind);
    }
}

/* Excitation list in the reference's loop order, for ket configurations
 * [q0, q1): one entry (Q, idx, P, sign) per surviving (Q, operator-index
 * tuple).  Writes at most cap entries, returns the total count (so cap = 0
 * counts).  idx is the flat row-major tensor index, leftmost operator most
 * significant. */
i64 orc_excitations(const char *ops, i64 n_orbs, i64 n_core, const i64 *comb_bra,
                    i64 n_elec_ket, const i64 *ket_cfgs, i64 q0, i64 q1,
                    i64 cap, i64 *e_q, i64 *e_idx, i64 *e_p, int8_t *e_sign)
{
    walker w;
    walker_init(&w, ops, n_orbs, n_core, n_elec_ket, comb_bra);
    w.mode = (cap > 0) ? 1 : 2;
    w.cap = cap;
    w.e_q = e_q;
    w.e_idx = e_idx;
    w.e_p = e_p;
    w.e_sign = e_sign;
    i64 ind[8];
    for (i64 Q = q0; Q < q1; Q++) {
        w.Q = Q;
        walk(&w, (int)w.k - 1, ket_cfgs + Q * n_elec_ket, n_elec_ket, 1, ind);
    }
    return w.n_emit;
}

/* ------------------------------------------------------------------------ */
/* scalar contractions (general-XRCC/H_contractions.c)                      */
/* ------------------------------------------------------------------------ */

#define IX2(a, b, n) ((a) * (n) + (b))

/* sum h[p,q] R[p,q]  (monomer_1e :48 and monomer_extPot :80) */
double orc_monomer_1e(i64 n, const double *Rca, const double *h)
{
    double H = 0.0;
    for (i64 p = 0; p < n; p++)
        for (i64 q = 0; q < n; q++) H += h[IX2(p, q, n)] * Rca[IX2(p, q, n)];
    return H;
}

/* sum V[p,q,r,s] R[p,q,s,r]  (monomer_2e :61) */
double orc_monomer_2e(i64 n, const double *Rccaa, const double *V)
{
    double H = 0.0;
    for (i64 p = 0; p < n; p++)
        for (i64 q = 0; q < n; q++)
            for (i64 r = 0; r < n; r++)
                for (i64 s = 0; s < n; s++)
                    H += V[((p * n + q) * n + r) * n + s] * Rccaa[((p * n + q) * n + s) * n + r];
    return H;
}

/* sum V[p1,q1,r2,s2] Rcc1[p1,q1] Raa2[s2,r2]  (dimer_2min2pls :95) */
double orc_dimer_2min2pls(i64 n1, i64 n2, const double *Rcc1, const double *Raa2,
                          const double *V)
{
    double H = 0.0;
    for (i64 p = 0; p < n1; p++)
        for (i64 q = 0; q < n1; q++) {
            double a = Rcc1[IX2(p, q, n1)];
            for (i64 r = 0; r < n2; r++)
                for (i64 s = 0; s < n2; s++)
                    H += V[((p * n1 + q) * n2 + r) * n2 + s] * a * Raa2[IX2(s, r, n2)];
        }
    return H;
}

/* sum h[p1,q2] Rc1[p1] Ra2[q2]  (dimer_1min1pls_1e :116) */
double orc_dimer_1min1pls_1e(i64 n1, i64 n2, const double *Rc1, const double *Ra2,
                             const double *h)
{
    double H = 0.0;
    for (i64 p = 0; p < n1; p++)
        for (i64 q = 0; q < n2; q++) H += h[IX2(p, q, n2)] * Rc1[p] * Ra2[q];
    return H;
}

/* 2*( sum V1112[p1,q1,r1,s2] Rcca1[q1,p1,r1] Ra2[s2]
 *   + sum V1222[p1,q2,r2,s2] Rc1[p1] Rcaa2[q2,s2,r2] )  (dimer_1min1pls_2e :129) */
double orc_dimer_1min1pls_2e(i64 n1, i64 n2, const double *Rc1, const double *Rcca1,
                             const double *Ra2, const double *Rcaa2,
                             const double *V1112, const double *V1222)
{
    double H = 0.0;
    for (i64 p = 0; p < n1; p++)
        for (i64 q = 0; q < n1; q++)
            for (i64 r = 0; r < n1; r++) {
                double a = Rcca1[(q * n1 + p) * n1 + r];
                for (i64 s = 0; s < n2; s++)
                    H += V1112[((p * n1 + q) * n1 + r) * n2 + s] * a * Ra2[s];
            }
    for (i64 p = 0; p < n1; p++)
        for (i64 q = 0; q < n2; q++)
            for (i64 r = 0; r < n2; r++)
                for (i64 s = 0; s < n2; s++)
                    H += V1222[((p * n2 + q) * n2 + r) * n2 + s] * Rc1[p] * Rcaa2[(q * n2 + s) * n2 + r];
    return 2.0 * H;
}

/* 4 * sum V[p1,q2,r1,s2] Rca1[p1,r1] Rca2[q2,s2]  (dimer_ExEx :163) */
double orc_dimer_ExEx(i64 n1, i64 n2, const double *Rca1, const double *Rca2,
                      const double *V)
{
    double H = 0.0;
    for (i64 p = 0; p < n1; p++)
        for (i64 q = 0; q < n2; q++)
            for (i64 r = 0; r < n1; r++) {
                double a = Rca1[IX2(p, r, n1)];
                for (i64 s = 0; s < n2; s++)
                    H += V[((p * n2 + q) * n1 + r) * n2 + s] * a * Rca2[IX2(q, s, n2)];
            }
    return 4.0 * H;
}

/* 2 * sum V[p1,q1,r2,s3] Rcc1[q1,p1] Ra2[r2] Ra3[s3]  (trimer_2min1pls1pls :184) */
double orc_trimer_2min1pls1pls(i64 n1, i64 n2, i64 n3, const double *Rcc1,
                               const double *Ra2, const double *Ra3, const double *V)
{
    double H = 0.0;
    for (i64 p = 0; p < n1; p++)
        for (i64 q = 0; q < n1; q++)
            for (i64 r = 0; r < n2; r++)
                for (i64 s = 0; s < n3; s++)
                    H += V[((p * n1 + q) * n2 + r) * n3 + s] * Rcc1[IX2(q, p, n1)] * Ra2[r] * Ra3[s];
    return 2.0 * H;
}

/* 2 * sum V[r2,s3,p1,q1] Raa1[q1,p1] Rc2[r2] Rc3[s3]  (trimer_2pls1min1min :208) */
double orc_trimer_2pls1min1min(i64 n1, i64 n2, i64 n3, const double *Raa1,
                               const double *Rc2, const double *Rc3, const double *V)
{
    double H = 0.0;
    for (i64 r = 0; r < n2; r++)
        for (i64 s = 0; s < n3; s++)
            for (i64 p = 0; p < n1; p++)
                for (i64 q = 0; q < n1; q++)
                    H += V[((r * n3 + s) * n1 + p) * n1 + q] * Rc2[r] * Rc3[s] * Raa1[IX2(q, p, n1)];
    return 2.0 * H;
}

/* 4 * sum V[p1,r2,q1,s3] Rca1[p1,q1] Rc2[r2] Ra3[s3]  (trimer_Ex1min1pls :232) */
double orc_trimer_Ex1min1pls(i64 n1, i64 n2, i64 n3, const double *Rca1,
                             const double *Rc2, const double *Ra3, const double *V)
{
    double H = 0.0;
    for (i64 p = 0; p < n1; p++)
        for (i64 r = 0; r < n2; r++)
            for (i64 q = 0; q < n1; q++)
                for (i64 s = 0; s < n3; s++)
                    H += V[((p * n2 + r) * n1 + q) * n3 + s] * Rc2[r] * Rca1[IX2(p, q, n1)] * Ra3[s];
    return 4.0 * H;
}
