// LASPack-compatible C-ABI (include/laspack.h) with the device Krylov
// backend. Host-side QMatrix rows are built by Q_SetLen / Q_SetEntry exactly
// as LASPack's callers expect (src/navier_velocity.cpp:340-345 and the other
// call sites); the first solve after a change converts them to a device ELL
// matrix (symmetric storage expanded), the vectors cross PCIe, and x comes
// back in Cmp[1..Dim]. The device-resident path (cfd_solver.h) never takes
// this route; it is the compatibility boundary for unmodified callers.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "krylov.h"
#include "laspack.h"

namespace cfd {
namespace {

// Device matrix of one sparsity structure. The reference builds a new
// QMatrix with the same pattern every outer iteration
// (src/navier_pressure.cpp:14-18,74), so the ELL columns, the colouring and
// each entry's ELL position are kept per structure; a new QMatrix with a
// known pattern only sends its values (one pinned copy and a scatter kernel).
struct DevQ {
    int n = 0, K = 0;
    bool sym = false;
    size_t nent = 0;
    std::vector<size_t> rowptr;  // host: first flat entry of each row
    std::vector<int> cols;       // host: 0-based column of each flat entry (the key)
    DBuf<double> diag, val, flat;
    DBuf<int> col, own, mir;
    DBuf<size_t> rp;
    const QMatrix *owner = nullptr;  // the QMatrix whose values the device holds
    std::vector<int> hcol;
    double *staging = nullptr;   // pinned host copy of the flat values
    Colouring colour;
    bool colour_ok = false;
    Ell ell() { return Ell{n, K, diag.p, val.p, col.p}; }
    ~DevQ() {
        if (staging) (void)hipHostFree(staging);
    }
};

struct LasState {
    double eps = 1e-8;
    int last_iter = 0;
    double last_acc = 0.;
    LASErrIdType err = LASOK;
    int gmres_m = 10;
    std::unique_ptr<Krylov> kry;
    DBuf<double> x, b;
    std::vector<std::unique_ptr<DevQ>> structs;  // most recently used last
    Krylov &K() {
        if (!kry) {
            hipStream_t s;
            CFD_HIP(hipStreamCreate(&s));
            kry.reset(new Krylov(s));
        }
        return *kry;
    }
};
LasState &las() {
    static LasState st;
    return st;
}

void set_err(LASErrIdType e) {
    if (las().err == LASOK) las().err = e;
}

// values of the flat entries -> ELL: diagonal entries summed in entry order,
// each off-diagonal entry to its slot (and, for symmetric storage, to the
// mirrored slot of the other row); every slot is written by one entry
__global__ void k_scatter_entries(int n, const size_t *__restrict__ rp,
                                  const int *__restrict__ own, const int *__restrict__ mir,
                                  const double *__restrict__ flat, double *__restrict__ diag,
                                  double *__restrict__ val) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double d = 0.;
        for (size_t e = rp[r]; e < rp[r + 1]; ++e) {
            const double v = flat[e];
            const int o = own[e];
            if (o < 0) {
                d += v;
                continue;
            }
            val[o] = v;
            if (mir[e] >= 0) val[mir[e]] = v;
        }
        diag[r] = d;
    }
}

// f(r0, r1) over the 1-based rows [1, n], in up to 16 host threads
template <class F>
bool rows_parallel(int n, F f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int T = n < (1 << 16) ? 1 : (int)std::min(16u, hw);
    std::vector<char> ok(T, 1);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int r0 = 1 + (int)((long long)n * t / T), r1 = 1 + (int)((long long)n * (t + 1) / T);
        if (t + 1 == T) ok[t] = f(r0, r1);
        else th.emplace_back([&ok, &f, t, r0, r1] { ok[t] = f(r0, r1); });
    }
    for (auto &x : th) x.join();
    for (char o : ok)
        if (!o) return false;
    return true;
}

// does Q have the pattern of d? (row lengths and columns, entry by entry)
bool same_structure(const QMatrix *Q, const DevQ &d) {
    const int n = (int)Q->Dim;
    if (d.n != n || d.sym != (Q->Symmetry == True)) return false;
    return rows_parallel(n, [&](int r0, int r1) {
        for (int r = r0; r < r1; ++r) {
            const size_t e0 = d.rowptr[r - 1];
            if (Q->Len[r] != d.rowptr[r] - e0) return false;
            const size_t *c = Q->Col[r];
            for (size_t e = 0; e < Q->Len[r]; ++e)
                if ((int)c[e] - 1 != d.cols[e0 + e]) return false;
        }
        return true;
    });
}

// the structure of Q: flat entries, ELL columns (slot order: a row's
// off-diagonals as they are met scanning the rows in order, mirrored entries
// of symmetric storage included), each entry's slot, the device arrays
std::unique_ptr<DevQ> build_structure(const QMatrix *Q, hipStream_t s) {
    std::unique_ptr<DevQ> d(new DevQ);
    const int n = (int)Q->Dim;
    d->n = n;
    d->sym = Q->Symmetry == True;
    d->rowptr.assign(n + 1, 0);
    for (int r = 1; r <= n; ++r) d->rowptr[r] = d->rowptr[r - 1] + Q->Len[r];
    d->nent = d->rowptr[n];
    if (d->nent >= (size_t)1 << 31) throw HipError("more than 2^31 matrix entries");
    d->cols.resize(d->nent);
    std::vector<int> cnt(n, 0);
    for (int r = 1; r <= n; ++r)
        for (size_t e = 0; e < Q->Len[r]; ++e) {
            const size_t c = Q->Col[r][e];
            if (c < 1 || c > (size_t)n) throw HipError("Q_SetEntry column out of range");
            d->cols[d->rowptr[r - 1] + e] = (int)c - 1;
            if ((int)c != r) {
                cnt[r - 1]++;
                if (d->sym) cnt[c - 1]++;
            }
        }
    int K = 1;
    for (int c : cnt) K = std::max(K, c);
    if (K > 64) throw HipError("more than 64 off-diagonal entries in a row");
    d->K = K;
    d->hcol.resize((size_t)K * n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < K; ++k) d->hcol[(size_t)k * n + i] = i;  // padding: the row itself
    std::vector<int> own(d->nent, -1), mir(d->nent, -1), fill(n, 0);
    for (int r = 0; r < n; ++r)
        for (size_t e = d->rowptr[r]; e < d->rowptr[r + 1]; ++e) {
            const int c = d->cols[e];
            if (c == r) continue;
            const int k = fill[r]++;
            own[e] = k * n + r;
            d->hcol[(size_t)k * n + r] = c;
            if (d->sym) {
                const int kc = fill[c]++;
                mir[e] = kc * n + c;
                d->hcol[(size_t)kc * n + c] = r;
            }
        }
    d->diag.alloc(n);
    d->val.alloc((size_t)K * n);
    d->val.zero(s);  // padding slots stay zero
    d->col.alloc((size_t)K * n);
    d->col.upload(d->hcol.data(), (size_t)K * n, s);
    d->rp.alloc(n + 1);
    d->rp.upload(d->rowptr.data(), n + 1, s);
    d->own.alloc(std::max<size_t>(d->nent, 1));
    d->mir.alloc(std::max<size_t>(d->nent, 1));
    d->own.upload(own.data(), d->nent, s);
    d->mir.upload(mir.data(), d->nent, s);
    d->flat.alloc(std::max<size_t>(d->nent, 1));
    CFD_HIP(hipHostMalloc((void **)&d->staging, sizeof(double) * std::max<size_t>(d->nent, 1),
                          hipHostMallocDefault));
    return d;
}

DevQ *upload(QMatrix *Q) {
    LasState &L = las();
    hipStream_t s = L.K().stream();
    DevQ *d = nullptr;
    for (size_t i = 0; i < L.structs.size(); ++i)
        if (L.structs[i].get() == (DevQ *)Q->Dev) d = L.structs[i].get();
    if (d && !Q->Dirty && d->owner == Q) return d;
    d = nullptr;
    for (size_t i = 0; i < L.structs.size(); ++i)
        if (same_structure(Q, *L.structs[i])) {
            std::unique_ptr<DevQ> hit = std::move(L.structs[i]);  // most recently used last
            L.structs.erase(L.structs.begin() + i);
            L.structs.push_back(std::move(hit));
            d = L.structs.back().get();
            break;
        }
    if (!d) {
        if (L.structs.size() >= 8) L.structs.erase(L.structs.begin());
        L.structs.push_back(build_structure(Q, s));
        d = L.structs.back().get();
    }
    // the values: one pinned copy, one transfer, the scatter on the device
    const int n = d->n;
    rows_parallel(n, [&](int r0, int r1) {
        for (int r = r0; r < r1; ++r)
            std::memcpy(d->staging + d->rowptr[r - 1], Q->Val[r], sizeof(double) * Q->Len[r]);
        return true;
    });
    CFD_HIP(hipMemcpyAsync(d->flat.p, d->staging, sizeof(double) * d->nent,
                           hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scatter_entries, dim3(grid_for(n)), dim3(kBlock), 0, s, n,
                       (const size_t *)d->rp.p, (const int *)d->own.p, (const int *)d->mir.p,
                       (const double *)d->flat.p, d->diag.p, d->val.p);
    CFD_LAUNCH_CHECK();
    CFD_HIP(hipStreamSynchronize(s));  // the staging buffer is reused by the next upload
    Q->Dev = d;
    Q->Dirty = 0;
    d->owner = Q;
    return d;
}

int pc_kind(PrecondProcType p) {
    if (p == nullptr) return CFD_PC_NONE;
    if (p == JacobiPrecond) return CFD_PC_JACOBI;
    if (p == SSORPrecond) return CFD_PC_SSOR;
    if (p == ILUPrecond) return CFD_PC_ILU;
    return -1;
}

Vector *run(int method, QMatrix *A, Vector *x, Vector *b, int maxit, PrecondProcType pp,
            double omega) {
    LasState &L = las();
    if (L.err != LASOK) return x;
    if (!A || !x || !b || !x->Cmp || !b->Cmp) { set_err(LASLValErr); return x; }
    if (A->Dim != x->Dim || A->Dim != b->Dim) { set_err(LASDimErr); return x; }
    int pc = pc_kind(pp);
    if (pc < 0) { set_err(LASUserErr); return x; }
    try {
        DevQ *d = upload(A);
        Krylov &K = L.K();
        hipStream_t s = K.stream();
        const int n = d->n;
        L.x.ensure(n);
        L.b.ensure(n);
        L.x.upload(x->Cmp + 1, n, s);
        L.b.upload(b->Cmp + 1, n, s);
        Colouring *col = nullptr;
        if (pc == CFD_PC_SSOR || pc == CFD_PC_ILU) {
            if (!d->colour_ok) {
                d->colour.build(d->hcol.data(), n, d->K, nullptr, d->col.p, s);
                d->colour_ok = true;
            }
            col = &d->colour;
        }
        cfd_solve_cfg cfg{method, pc, omega, L.eps, maxit};
        SolveStats st = K.solve(d->ell(), L.b.p, L.x.p, cfg, col);
        L.x.download(x->Cmp + 1, n, s);
        CFD_HIP(hipStreamSynchronize(s));
        L.last_iter = st.iters;
        L.last_acc = st.relres;
        if (st.breakdown) set_err(LASBreakdownErr);
    } catch (const std::exception &e) {
        last_error() = e.what();
        set_err(LASDeviceErr);
    }
    return x;
}

Vector *apply(int pc, QMatrix *A, Vector *y, Vector *c, double omega) {
    LasState &L = las();
    if (L.err != LASOK) return y;
    if (!A || !y || !c || A->Dim != y->Dim || A->Dim != c->Dim) { set_err(LASDimErr); return y; }
    try {
        DevQ *d = upload(A);
        Krylov &K = L.K();
        hipStream_t s = K.stream();
        const int n = d->n;
        L.x.ensure(n);
        L.b.ensure(n);
        L.b.upload(c->Cmp + 1, n, s);
        Colouring *col = nullptr;
        if (pc == CFD_PC_SSOR || pc == CFD_PC_ILU) {
            if (!d->colour_ok) {
                d->colour.build(d->hcol.data(), n, d->K, nullptr, d->col.p, s);
                d->colour_ok = true;
            }
            col = &d->colour;
        }
        K.precond(d->ell(), pc, omega, L.b.p, L.x.p, col);
        L.x.download(y->Cmp + 1, n, s);
        CFD_HIP(hipStreamSynchronize(s));
    } catch (const std::exception &e) {
        last_error() = e.what();
        set_err(LASDeviceErr);
    }
    return y;
}

}  // namespace
}  // namespace cfd

using namespace cfd;

extern "C" {

void V_Constr(Vector *V, char *Name, size_t Dim, InstanceType Instance, Boolean OwnData) {
    std::memset(V, 0, sizeof(*V));
    V->Name = Name ? strdup(Name) : nullptr;
    V->Dim = Dim;
    V->Instance = Instance;
    V->Multipl = 1.0;
    V->OwnData = OwnData;
    if (OwnData) {
        V->Cmp = (double *)calloc(Dim + 1, sizeof(double));
        if (!V->Cmp) set_err(LASMemAllocErr);
    }
}

void V_Destr(Vector *V) {
    if (!V) return;
    free(V->Name);
    if (V->OwnData) free(V->Cmp);
    V->Name = nullptr;
    V->Cmp = nullptr;
    V->Dim = 0;
}

void V_SetAllCmp(Vector *V, double Val) {
    if (!V || !V->Cmp) { set_err(LASLValErr); return; }
    for (size_t i = 1; i <= V->Dim; ++i) V->Cmp[i] = Val;
}

size_t V_GetDim(Vector *V) { return V ? V->Dim : 0; }

void Q_Constr(QMatrix *Q, char *Name, size_t Dim, Boolean Symmetry, ElOrderType ElOrder,
              InstanceType Instance, Boolean OwnData) {
    std::memset(Q, 0, sizeof(*Q));
    Q->Name = Name ? strdup(Name) : nullptr;
    Q->Dim = Dim;
    Q->Symmetry = Symmetry;
    Q->ElOrder = ElOrder;
    Q->Instance = Instance;
    Q->OwnData = OwnData;
    Q->Len = (size_t *)calloc(Dim + 1, sizeof(size_t));
    Q->Cap = (size_t *)calloc(Dim + 1, sizeof(size_t));
    Q->Col = (size_t **)calloc(Dim + 1, sizeof(size_t *));
    Q->Val = (double **)calloc(Dim + 1, sizeof(double *));
    if (!Q->Len || !Q->Cap || !Q->Col || !Q->Val) set_err(LASMemAllocErr);
    Q->Dirty = 1;
}

void Q_Destr(QMatrix *Q) {
    if (!Q) return;
    for (size_t r = 0; Q->Col && r <= Q->Dim; ++r) {
        free(Q->Col[r]);
        free(Q->Val[r]);
    }
    free(Q->Len);
    free(Q->Cap);
    free(Q->Col);
    free(Q->Val);
    free(Q->Name);  // the device matrix belongs to the structure cache
    std::memset(Q, 0, sizeof(*Q));
}

void Q_SetLen(QMatrix *Q, size_t RoC, size_t Len) {
    if (!Q || RoC < 1 || RoC > Q->Dim) { set_err(LASRangeErr); return; }
    if (Len > Q->Cap[RoC]) {
        size_t *c = (size_t *)realloc(Q->Col[RoC], Len * sizeof(size_t));
        double *v = (double *)realloc(Q->Val[RoC], Len * sizeof(double));
        if (!c || !v) { set_err(LASMemAllocErr); return; }
        Q->Col[RoC] = c;
        Q->Val[RoC] = v;
        Q->Cap[RoC] = Len;
    }
    for (size_t e = 0; e < Len; ++e) {
        Q->Col[RoC][e] = RoC;
        Q->Val[RoC][e] = 0.;
    }
    Q->Len[RoC] = Len;
    Q->Dirty = 1;
}

void Q_SetEntry(QMatrix *Q, size_t RoC, size_t Entry, size_t Pos, double Val) {
    if (!Q || RoC < 1 || RoC > Q->Dim || Entry >= Q->Len[RoC] || Pos < 1 || Pos > Q->Dim) {
        set_err(LASRangeErr);
        return;
    }
    if (Q->Symmetry && Pos < RoC) { set_err(LASSymStorErr); return; }
    Q->Col[RoC][Entry] = Pos;
    Q->Val[RoC][Entry] = Val;
    Q->Dirty = 1;
}

size_t Q_GetDim(QMatrix *Q) { return Q ? Q->Dim : 0; }

Vector *CGIter(QMatrix *A, Vector *x, Vector *b, int MaxIter, PrecondProcType P, double w) {
    return run(CFD_CG, A, x, b, MaxIter, P, w);
}
Vector *BiCGSTABIter(QMatrix *A, Vector *x, Vector *b, int MaxIter, PrecondProcType P, double w) {
    return run(CFD_BICGSTAB, A, x, b, MaxIter, P, w);
}
Vector *GMRESIter(QMatrix *A, Vector *x, Vector *b, int MaxIter, PrecondProcType P, double w) {
    return run(CFD_GMRES, A, x, b, MaxIter, P, w);
}
void SetGMRESRestart(int MaxSteps) {
    if (MaxSteps != 10) set_err(LASRangeErr);  // the device GMRES is GMRES(10)
}

Vector *JacobiPrecond(QMatrix *A, Vector *y, Vector *c, double Omega) {
    return apply(CFD_PC_JACOBI, A, y, c, Omega);
}
Vector *SSORPrecond(QMatrix *A, Vector *y, Vector *c, double Omega) {
    return apply(CFD_PC_SSOR, A, y, c, Omega);
}
Vector *ILUPrecond(QMatrix *A, Vector *y, Vector *c, double Omega) {
    return apply(CFD_PC_ILU, A, y, c, Omega);
}

void SetRTCAccuracy(double Eps) { las().eps = Eps; }
int GetLastNoIter(void) { return las().last_iter; }
double GetLastAccuracy(void) { return las().last_acc; }
LASErrIdType LASResult(void) { return las().err; }
Boolean LASBreak(void) { return las().err != LASOK ? True : False; }
void LASClearResult(void) { las().err = LASOK; }

}  // extern "C"
