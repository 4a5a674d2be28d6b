/*
 * lbm_oracle_impl.h -- body of the CPU oracle, compiled once per precision
 * by lbm_oracle.c with REAL (double or float) and FN(name) defined.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Every expression keeps the reference's operand order and literal types
 * (double literals such as 3.0, 0.5, 1. promote float operands exactly as
 * the C++ reference does), so the float instance reproduces the reference
 * as shipped and the double instance the reference built with real=double.
 */

struct FN(oracle) {
    oracle_params p;
    const oracle_lattice *L;
    int Q;
    long N;
    REAL w[27];
    REAL cs, rho_lbm, u_lbm, omega;
    int *adj;          /* push targets, node-major [g*Q+s] */
    int *type;
    REAL *fEven, *fOdd, *fEq;
    REAL *ux, *uy, *uz, *rho;
    REAL *fOut;        /* buffer written by the last step */
    long ts;
    int closure;       /* learned closure in place of Smagorinsky */
    float mlp[1345];
    double cl_in, cl_out;
};

/* LatticeStructure.hpp:92-103 */
static void FN(equilibrium)(struct FN(oracle) *o, long nd)
{
    const oracle_lattice *L = o->L;
    REAL *feq = o->fEq + nd * o->Q;
    REAL cu;
    for (int s = 0; s < o->Q; ++s) {
        cu = 3.0 * (L->ex[s] * o->ux[nd] + L->ey[s] * o->uy[nd] + L->ez[s] * o->uz[nd]);
#if ORACLE_PLANT == 1
        feq[s] = o->w[s] * o->rho[nd] * (1.0 + cu -
#else
        feq[s] = o->w[s] * o->rho[nd] *
                 (1.0 + cu + 0.5 * (cu * cu) -
#endif
                  3. / 2. * (o->ux[nd] * o->ux[nd] + o->uy[nd] * o->uy[nd] + o->uz[nd] * o->uz[nd]));
    }
}

/* LatticeStructure.hpp:140-153 (sqrt is the double routine, as the
 * reference's unqualified call resolves to ::sqrt(double)). */
static REAL FN(turbulence)(REAL omega, const REAL *S, REAL cs)
{
    REAL omega_r, nu, nu_e, P;
    nu = ((1. / omega) - 0.5) / 3.0;
    P = sqrt(S[0] * S[0] + S[4] * S[4] + S[8] * S[8] +
             2.0 * (S[1] * S[1] + S[2] * S[2] + S[5] * S[5]));
    P *= cs;
#if ORACLE_PLANT == 5
    P *= cs;
#endif
    P = sqrt(P + nu * nu) - nu;
    nu_e = P / 6.;
    omega_r = 1. / (3. * (nu + nu_e) + 0.5);
    return omega_r;
}

/* TLBM_Partition.cpp:738-813 for one node, then stream_node_data 826-837 */
static void FN(process_node)(struct FN(oracle) *o, REAL *fOut, REAL *fIn, long nd)
{
    const oracle_lattice *L = o->L;
    const int Q = o->Q;
    REAL *f = fIn + nd * Q;
    REAL *feq = o->fEq + nd * Q;
    const int t = o->type[nd];

    /* compute_macroscopic_data, LatticeStructure.hpp:69-90 */
    {
        REAL ux_l = 0, uy_l = 0, uz_l = 0, rho_l = 0, fv;
        for (int s = 0; s < Q; ++s) {
            fv = f[s];
            rho_l += fv;
            ux_l += fv * L->ex[s];
            uy_l += fv * L->ey[s];
            uz_l += fv * L->ez[s];
        }
        ux_l /= rho_l;
        uy_l /= rho_l;
        uz_l /= rho_l;
        o->ux[nd] = ux_l; o->uy[nd] = uy_l; o->uz[nd] = uz_l; o->rho[nd] = rho_l;
    }

    if (t == 0)
        FN(equilibrium)(o, nd);
    if (t == 1) {
        /* solid: u = 0, equilibrium, whole-vector bounce-back
         * (TLBM_Partition.cpp:750-756, D3Q19LatticeStructure.hpp:65-76) */
        REAL tmp[27];
        o->ux[nd] = 0; o->uy[nd] = 0; o->uz[nd] = 0;
        FN(equilibrium)(o, nd);
        for (int s = 0; s < Q; ++s) tmp[s] = f[L->bb[s]];
        for (int s = 0; s < Q; ++s) f[s] = tmp[s];
    }
    if (t != 1) {
        if (t == 2) {
            /* set_inletW_bc_macro / _micro */
            const REAL u_bc = o->u_lbm;
            REAL s_in = 0, sz = 0;
            for (int i = 0; i < L->n_unk; ++i) s_in += f[L->in_list[i]];
            for (int i = 0; i < L->n_zero; ++i) sz += f[L->zero_list[i]];
            o->ux[nd] = 0; o->uy[nd] = 0; o->uz[nd] = u_bc;
            if (L->in_int_two)
                o->rho[nd] = (1. / (1. - u_bc)) * (2 * s_in + sz);
            else
                o->rho[nd] = (1. / (1. - u_bc)) * (2. * s_in + sz);
            FN(equilibrium)(o, nd);
            for (int i = 0; i < L->n_unk; ++i) {
#if ORACLE_PLANT == 3
                const int par = L->bb[L->in_unk[i]];
#else
                const int par = L->in_par[i];
#endif
                f[L->in_unk[i]] = feq[L->in_unk[i]] + f[par] - feq[par];
            }
        }
        if (t == 3) {
            /* set_outletE_bc_macro / _micro; ux, uy stay from the moments */
            const REAL rho_bc = o->rho_lbm;
            REAL sout = 0, sz = 0;
            o->rho[nd] = rho_bc;
            for (int i = 0; i < L->n_unk; ++i) sout += f[L->out_list[i]];
            for (int i = 0; i < L->n_zero; ++i) sz += f[L->zero_list[i]];
#if ORACLE_PLANT == 4
            o->uz[nd] = 1. - (1. / rho_bc) * (2. * sout + sz);
#else
            o->uz[nd] = -1. + (1. / rho_bc) * (2. * sout + sz);
#endif
            FN(equilibrium)(o, nd);
            for (int i = 0; i < L->n_unk; ++i)
                f[L->out_unk[i]] = feq[L->out_unk[i]] + f[L->out_par[i]] - feq[L->out_par[i]];
        }
        if (t == 5) {
            /* set_uz_bc, LatticeStructure.hpp:166-177: the y term reads
             * uy at node index s (not nd), as the reference does */
            const REAL u_bc = o->u_lbm;
            for (int s = 0; s < Q; ++s)
                f[s] += 3. * o->rho[nd] * o->w[s] *
                        (L->ez[s] * (u_bc - o->uz[nd]) + L->ex[s] * (0. - o->ux[nd]) +
#if ORACLE_PLANT == 6
                         L->ey[s] * (0. - o->uy[nd]));
#else
                         L->ey[s] * (0. - o->uy[s]));
#endif
            FN(equilibrium)(o, nd);
        }

        REAL omega = o->omega;
        if (o->closure) {
            /* learned closure: the stress of compute_strain_tensor and its
             * norm, scaled, through the network (bf16 inference) */
            double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int s = 0; s < Q; ++s) {
                const double e3[3] = {L->ex[s], L->ey[s], L->ez[s]};
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) S[i * 3 + j] += e3[i] * e3[j] * ((double)f[s] - (double)feq[s]);
            }
            const double qn = sqrt(S[0] * S[0] + S[4] * S[4] + S[8] * S[8] +
                                   2.0 * (S[1] * S[1] + S[2] * S[2] + S[5] * S[5]));
            const float x[7] = {(float)(S[0] * o->cl_in), (float)(S[4] * o->cl_in),
                                (float)(S[8] * o->cl_in), (float)(S[1] * o->cl_in),
                                (float)(S[2] * o->cl_in), (float)(S[5] * o->cl_in),
                                (float)(qn * o->cl_in)};
            double z;
            mlp_infer_bf16(o->mlp, x, 1, &z);
            const double sp = z > 30.0 ? z : log1p(exp(z));
            omega = (REAL)(1.0 / (1.0 / o->p.omega + o->cl_out * sp));
        } else if (o->cs > 0) {
            /* compute_strain_tensor, LatticeStructure.hpp:105-122 */
            REAL S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            REAL e[3];
            for (int s = 0; s < Q; ++s) {
                e[0] = L->ex[s]; e[1] = L->ey[s]; e[2] = L->ez[s];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        S[i * 3 + j] += e[i] * e[j] * (f[s] - feq[s]);
            }
            omega = FN(turbulence)(omega, S, o->cs);
        }
        if (o->p.dynamics == 1) {
            /* relax, LatticeStructure.hpp:155-163 */
            for (int s = 0; s < Q; ++s) f[s] = f[s] - omega * (f[s] - feq[s]);
        }
    }

    /* stream_node_data: push to the periodic neighbour */
    for (int s = 0; s < Q; ++s) {
        const long t_nd = o->adj[nd * Q + s];
        fOut[t_nd * Q + s] = f[s];
    }
}

struct FN(oracle) *FN(oracle_create)(const oracle_params *p, const int *ndType)
{
    const oracle_lattice *L = oracle_lattice_for(p->lattice);
    if (!L || p->nx <= 0 || p->ny <= 0 || p->nz <= 0 || !ndType) return 0;
    const long N = (long)p->nx * p->ny * p->nz;
    for (long g = 0; g < N; ++g) {
        const int t = ndType[g];
        if (!(t == 0 || t == 1 || t == 2 || t == 3 || t == 5)) return 0;
    }
    struct FN(oracle) *o = (struct FN(oracle) *)calloc(1, sizeof(*o));
    if (!o) return 0;
    o->p = *p;
    o->L = L;
    o->Q = L->Q;
    o->N = N;
    for (int s = 0; s < L->Q; ++s)
        o->w[s] = (L->w_is_float && !p->exact_weights) ? (REAL)L->wf[s] : (REAL)L->wd[s];
    o->cs = (REAL)p->cs;
    o->rho_lbm = (REAL)p->rho_lbm;
    o->u_lbm = (REAL)p->u_lbm;
    o->omega = (REAL)p->omega;
    const int Q = L->Q;
    o->adj = (int *)malloc(sizeof(int) * N * Q);
    o->type = (int *)malloc(sizeof(int) * N);
    o->fEven = (REAL *)malloc(sizeof(REAL) * N * Q);
    o->fOdd = (REAL *)malloc(sizeof(REAL) * N * Q);
    o->fEq = (REAL *)calloc((size_t)N * Q, sizeof(REAL));
    o->ux = (REAL *)calloc(N, sizeof(REAL));
    o->uy = (REAL *)calloc(N, sizeof(REAL));
    o->uz = (REAL *)calloc(N, sizeof(REAL));
    o->rho = (REAL *)calloc(N, sizeof(REAL));
    if (!o->adj || !o->type || !o->fEven || !o->fOdd || !o->fEq || !o->ux || !o->uy || !o->uz ||
        !o->rho) {
        FN(oracle_destroy)(o);
        return 0;
    }
    memcpy(o->type, ndType, sizeof(int) * N);
    /* create_adj_matrix + get_tgt_index, TLBM_Partition.cpp:254-271, 691-737 */
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    for (long g = 0; g < N; ++g) {
        const int z = (int)(g / ((long)nx * ny));
        const int y = (int)((g - (long)z * nx * ny) / nx);
        const int x = (int)(g - (long)z * nx * ny - (long)y * nx);
        for (int s = 0; s < Q; ++s) {
#if ORACLE_PLANT == 2
            int tx = (x + L->ex[s]) % nx, ty = (y - L->ey[s]) % ny, tz = (z + L->ez[s]) % nz;
#else
            int tx = (x + L->ex[s]) % nx, ty = (y + L->ey[s]) % ny, tz = (z + L->ez[s]) % nz;
#endif
            if (tx < 0) tx += nx;
            if (ty < 0) ty += ny;
            if (tz < 0) tz += nz;
            o->adj[g * Q + s] = (int)(tx + (long)ty * nx + (long)tz * nx * ny);
        }
    }
    /* initialize_data_arrays, TLBM_Partition.cpp:448-479: f = w*rho */
    const REAL rho0 = o->rho_lbm;
    for (long g = 0; g < N; ++g)
        for (int s = 0; s < Q; ++s) {
            o->fEven[g * Q + s] = o->w[s] * rho0;
            o->fOdd[g * Q + s] = o->w[s] * rho0;
        }
    o->fOut = 0;
    o->ts = 0;
    return o;
}

void FN(oracle_destroy)(struct FN(oracle) *o)
{
    if (!o) return;
    free(o->adj); free(o->type); free(o->fEven); free(o->fOdd); free(o->fEq);
    free(o->ux); free(o->uy); free(o->uz); free(o->rho);
    free(o);
}

/* take_LBM_time_step, TLBM_Partition.cpp:910-947, driven as tLBM.cpp:53
 * does with isEven = ts % 2 */
void FN(oracle_step)(struct FN(oracle) *o, int nsteps)
{
    for (int k = 0; k < nsteps; ++k) {
        const int isEven = (o->ts % 2) != 0;
        REAL *fIn = isEven ? o->fEven : o->fOdd;
        REAL *fOut = isEven ? o->fOdd : o->fEven;
        for (long nd = 0; nd < o->N; ++nd) FN(process_node)(o, fOut, fIn, nd);
        o->fOut = fOut;
        ++o->ts;
    }
}

void FN(oracle_get_macros)(const struct FN(oracle) *o, REAL *ux, REAL *uy, REAL *uz, REAL *rho)
{
    memcpy(ux, o->ux, sizeof(REAL) * o->N);
    memcpy(uy, o->uy, sizeof(REAL) * o->N);
    memcpy(uz, o->uz, sizeof(REAL) * o->N);
    memcpy(rho, o->rho, sizeof(REAL) * o->N);
}

/* make_force_calc_map + calc_force, TLBM_Partition.cpp:143-208, on the
 * buffer the last step wrote.  Solid nodes with a type-0 neighbour only. */
void FN(oracle_get_forces)(const struct FN(oracle) *o, REAL *Fx, REAL *Fy, REAL *Fz)
{
    const oracle_lattice *L = o->L;
    const int Q = o->Q;
    const REAL *fOut = o->fOut ? o->fOut : o->fOdd;
    for (long nd = 0; nd < o->N; ++nd) {
        Fx[nd] = 0; Fy[nd] = 0; Fz[nd] = 0;
    }
    for (long nd = 0; nd < o->N; ++nd) {
        if (o->type[nd] != 1) continue;
        for (int spd = 0; spd < Q; ++spd) {
            const long ngb = o->adj[nd * Q + spd];
            if (o->type[ngb] != 0) continue;
            const int bbs = L->bb[spd];
            const REAL f1 = fOut[nd * Q + spd];
            const REAL f2 = fOut[ngb * Q + bbs];
            Fx[nd] += L->ex[bbs] * (f1 + f2);
            Fy[nd] += L->ey[bbs] * (f1 + f2);
            Fz[nd] += L->ez[bbs] * (f1 + f2);
        }
    }
}

void FN(oracle_get_populations)(const struct FN(oracle) *o, REAL *f)
{
    const REAL *src = o->fOut ? o->fOut : o->fOdd;
    memcpy(f, src, sizeof(REAL) * o->N * o->Q);
}

void FN(oracle_set_populations)(struct FN(oracle) *o, const REAL *f)
{
    memcpy(o->fEven, f, sizeof(REAL) * o->N * o->Q);
    memcpy(o->fOdd, f, sizeof(REAL) * o->N * o->Q);
    o->fOut = (o->ts % 2) != 0 ? o->fEven : o->fOdd;
}

void FN(oracle_set_closure)(struct FN(oracle) *o, const float *params, double cl_in, double cl_out)
{
    o->closure = params != 0;
    if (params) memcpy(o->mlp, params, sizeof(o->mlp));
    o->cl_in = cl_in;
    o->cl_out = cl_out;
}
