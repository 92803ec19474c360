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
#include <vector>

#include "krylov.h"
#include "laspack.h"

namespace cfd {
namespace {

struct DevQ {
    int n = 0, K = 0;
    DBuf<double> diag, val;
    DBuf<int> col;
    std::vector<int> hcol;
    Colouring colour;
    bool colour_ok = false;
    Ell ell() { return Ell{n, K, diag.p, val.p, col.p}; }
};

struct LasState {
    double eps = 1e-8;
    int last_iter = 0;
    double last_acc = 0.;
    LASErrIdType err = LASOK;
    int gmres_m = 10;
    std::unique_ptr<Krylov> kry;
    DBuf<double> x, b;
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

// rows -> device ELL (diagonal separate; duplicate diagonal entries summed)
DevQ *upload(QMatrix *Q) {
    DevQ *d = (DevQ *)Q->Dev;
    if (d && !Q->Dirty) return d;
    if (!d) {
        d = new DevQ;
        Q->Dev = d;
    }
    const int n = (int)Q->Dim;
    std::vector<double> diag(n, 0.);
    std::vector<std::vector<std::pair<int, double>>> off(n);
    for (int r = 1; r <= n; ++r)
        for (size_t e = 0; e < Q->Len[r]; ++e) {
            size_t c = Q->Col[r][e];
            if (c < 1 || c > (size_t)n) throw HipError("Q_SetEntry column out of range");
            double v = Q->Val[r][e];
            if ((int)c == r) diag[r - 1] += v;
            else {
                off[r - 1].push_back({(int)c - 1, v});
                if (Q->Symmetry) off[c - 1].push_back({r - 1, v});
            }
        }
    int K = 1;
    for (auto &o : off) K = std::max<int>(K, (int)o.size());
    if (K > 64) throw HipError("more than 64 off-diagonal entries in a row");
    std::vector<double> val((size_t)K * n, 0.);
    d->hcol.assign((size_t)K * n, 0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < K; ++k) {
            bool has = k < (int)off[i].size();
            d->hcol[(size_t)k * n + i] = has ? off[i][k].first : i;
            val[(size_t)k * n + i] = has ? off[i][k].second : 0.;
        }
    hipStream_t s = las().K().stream();
    d->n = n;
    d->K = K;
    d->diag.alloc(n);
    d->val.alloc((size_t)K * n);
    d->col.alloc((size_t)K * n);
    d->diag.upload(diag.data(), n, s);
    d->val.upload(val.data(), (size_t)K * n, s);
    d->col.upload(d->hcol.data(), (size_t)K * n, s);
    d->colour_ok = false;
    Q->Dirty = 0;
    CFD_HIP(hipStreamSynchronize(s));
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
    free(Q->Name);
    delete (DevQ *)Q->Dev;
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
