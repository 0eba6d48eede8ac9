/* Sanitizer driver of the C oracle (TEST INFRASTRUCTURE ONLY): runs every
 * density string (1 to 5 operators), the excitation lists and every
 * contraction of xr_oracle.c on small fragments (with and without a frozen
 * core, every electron count), with deterministic pseudo-random inputs, and
 * prints one checksum.  Built twice by oracle/Makefile -- plain and with
 * -fsanitize=address,undefined -- and run by tests/test_oracle_sanitize.py,
 * which requires a clean sanitizer run with the same checksum. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

void orc_combo_table(i64 n_e, i64 n_so, i64 *out);
void orc_density(const char *ops, double *out, i64 n_orbs, i64 n_core, i64 n_bra_states, i64 n_bra_cfg,
                 const double *zbra, const i64 *comb_bra, i64 n_ket_states, i64 n_ket_cfg, const double *zket,
                 i64 n_elec_ket, const i64 *ket_cfgs);
i64 orc_excitations(const char *ops, i64 n_orbs, i64 n_core, const i64 *comb_bra, i64 n_elec_ket,
                    const i64 *ket_cfgs, i64 q0, i64 q1, i64 cap, i64 *e_q, i64 *e_idx, i64 *e_p,
                    int8_t *e_sign);
double orc_monomer_1e(i64 n, const double *Rca, const double *h);
double orc_monomer_2e(i64 n, const double *Rccaa, const double *V);
double orc_dimer_2min2pls(i64 n1, i64 n2, const double *Rcc1, const double *Raa2, const double *V);
double orc_dimer_1min1pls_1e(i64 n1, i64 n2, const double *Rc1, const double *Ra2, const double *h);
double orc_dimer_1min1pls_2e(i64 n1, i64 n2, const double *Rc1, const double *Rcca1, const double *Ra2,
                             const double *Rcaa2, const double *V1112, const double *V1222);
double orc_dimer_ExEx(i64 n1, i64 n2, const double *Rca1, const double *Rca2, const double *V);

static uint64_t rng = 88172645463325252ull;
static double frand(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) / 9007199254740992.0 - 0.5;
}

/* every determinant of n_elec electrons with the n_core lowest spatial
 * orbitals of each spin filled, lexical order over the valence */
static i64 configs(i64 n_orbs, i64 n_core, i64 n_elec, i64 **out)
{
    i64 nv = 2 * (n_orbs - n_core), ne = n_elec - 2 * n_core;
    if (ne < 0 || ne > nv) { *out = NULL; return 0; }
    i64 val[64], cnt = 0, cap = 1;
    for (i64 i = 0; i < ne; i++) cap = cap * (nv - i) / (i + 1);
    *out = (i64 *)malloc((size_t)(cap * (n_elec > 0 ? n_elec : 1)) * sizeof(i64));
    for (i64 i = 0; i < ne; i++) val[i] = i;
    for (;;) {
        i64 full[128], k = 0;
        /* valence index v -> spin orbital: alpha v + n_core, beta v - (n_orbs - n_core) + n_orbs + n_core */
        for (i64 c = 0; c < n_core; c++) full[k++] = c;
        for (i64 i = 0; i < ne; i++) if (val[i] < n_orbs - n_core) full[k++] = val[i] + n_core;
        for (i64 c = 0; c < n_core; c++) full[k++] = n_orbs + c;
        for (i64 i = 0; i < ne; i++) if (val[i] >= n_orbs - n_core) full[k++] = val[i] + 2 * n_core;
        memcpy(*out + cnt * n_elec, full, (size_t)n_elec * sizeof(i64));
        cnt++;
        i64 i = ne - 1;
        while (i >= 0 && val[i] == nv - ne + i) i--;
        if (i < 0) break;
        val[i]++;
        for (i64 j = i + 1; j < ne; j++) val[j] = val[j - 1] + 1;
    }
    return cnt;
}

static const char *OPS[] = {"a", "c", "ca", "aa", "cc", "caa", "cca", "ccaa", "caaa", "ccca", "ccaaa", "cccaa"};

int main(void)
{
    double check = 0.0;
    const i64 cases[2][2] = {{3, 0}, {4, 1}};
    for (int cs = 0; cs < 2; cs++) {
        i64 n_orbs = cases[cs][0], n_core = cases[cs][1], dim = 2 * n_orbs;
        for (i64 nk = 2 * n_core; nk <= dim; nk++) {
            i64 *kc;
            i64 nQ = configs(n_orbs, n_core, nk, &kc);
            if (!nQ) continue;
            for (size_t o = 0; o < sizeof(OPS) / sizeof(OPS[0]); o++) {
                const char *ops = OPS[o];
                i64 k = (i64)strlen(ops), nb = nk;
                for (i64 l = 0; l < k; l++) nb += ops[l] == 'c' ? 1 : -1;
                i64 *bc;
                i64 nP = configs(n_orbs, n_core, nb, &bc);
                if (!nP) continue;
                i64 nvb = nb - 2 * n_core, nvso = dim - 2 * n_core;
                i64 *comb = (i64 *)calloc((size_t)((nvb > 0 ? nvb : 1) * nvso), sizeof(i64));
                if (nvb > 0) orc_combo_table(nvb, nvso, comb);
                const i64 nI = 2, nJ = 3;
                double *zb = (double *)malloc((size_t)(nI * nP) * sizeof(double));
                double *zk = (double *)malloc((size_t)(nJ * nQ) * sizeof(double));
                for (i64 i = 0; i < nI * nP; i++) zb[i] = frand();
                for (i64 i = 0; i < nJ * nQ; i++) zk[i] = frand();
                i64 ts = 1;
                for (i64 l = 0; l < k; l++) ts *= dim;
                double *out = (double *)calloc((size_t)(nI * nJ * ts), sizeof(double));
                orc_density(ops, out, n_orbs, n_core, nI, nP, zb, comb, nJ, nQ, zk, nk, kc);
                for (i64 i = 0; i < nI * nJ * ts; i++) check += out[i] * (double)((i % 7) + 1);
                i64 n = orc_excitations(ops, n_orbs, n_core, comb, nk, kc, 0, nQ, 0, NULL, NULL, NULL, NULL);
                i64 *eq = (i64 *)malloc((size_t)(n + 1) * sizeof(i64)), *ei = (i64 *)malloc((size_t)(n + 1) * sizeof(i64));
                i64 *ep = (i64 *)malloc((size_t)(n + 1) * sizeof(i64));
                int8_t *es = (int8_t *)malloc((size_t)(n + 1));
                orc_excitations(ops, n_orbs, n_core, comb, nk, kc, 0, nQ, n, eq, ei, ep, es);
                for (i64 i = 0; i < n; i++) check += 1e-3 * (double)(eq[i] + ei[i] + ep[i] + es[i]);
                free(eq); free(ei); free(ep); free(es);
                free(out); free(zb); free(zk); free(comb); free(bc);
            }
            free(kc);
        }
        /* contractions on random tensors of this fragment size */
        i64 n1 = dim, n2 = dim - 2;
        size_t big = (size_t)(n1 * n1 * n1 * n1 > n1 * n2 * n2 * n2 ? n1 * n1 * n1 * n1 : n1 * n2 * n2 * n2);
        double *R1 = malloc(big * sizeof(double)), *R2 = malloc(big * sizeof(double));
        double *R3 = malloc(big * sizeof(double)), *R4 = malloc(big * sizeof(double));
        double *V = malloc(big * sizeof(double)), *V2 = malloc(big * sizeof(double)), *h = malloc(big * sizeof(double));
        for (size_t i = 0; i < big; i++) {
            R1[i] = frand(); R2[i] = frand(); R3[i] = frand(); R4[i] = frand();
            V[i] = frand(); V2[i] = frand(); h[i] = frand();
        }
        check += orc_monomer_1e(n1, R1, h) + orc_monomer_2e(n1, R1, V);
        check += orc_dimer_2min2pls(n1, n2, R1, R2, V) + orc_dimer_ExEx(n1, n2, R1, R2, V);
        check += orc_dimer_1min1pls_1e(n1, n2, R1, R2, h);
        check += orc_dimer_1min1pls_2e(n1, n2, R1, R2, R3, R4, V, V2);
        free(R1); free(R2); free(R3); free(R4); free(V); free(V2); free(h);
    }
    printf("%.17g\n", check);
    return 0;
}
