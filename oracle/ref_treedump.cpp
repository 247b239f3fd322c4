// TEST INFRASTRUCTURE ONLY.  A thin extern "C" driver compiled against the
// reference headers and linked with the reference's own Cell/Field/KMeans
// objects (built from /root/reference/src by oracle/Makefile.ref).  It builds a
// reference Field and dumps its cells in pre-order, so the oracle's and the
// device's tree builds can be checked node by node against the literal
// reference (src/Field.cpp:163-277, src/Cell.cpp:556-615).  The reference
// sources are compiled in place; none is copied.
#include <complex>
#include <vector>
// The reference's Field.cpp is compiled here, where it lies, as part of this
// translation unit so every Field<D,C> needed below can be instantiated.
#include "Field.cpp"

template class Field<NData,Flat>; template class Field<NData,ThreeD>; template class Field<NData,Sphere>;
template class Field<KData,Flat>; template class Field<KData,ThreeD>; template class Field<KData,Sphere>;
template class Field<GData,Flat>; template class Field<GData,ThreeD>; template class Field<GData,Sphere>;
template class Field<ZData,Flat>; template class Field<ZData,ThreeD>; template class Field<ZData,Sphere>;
template class Field<VData,Flat>; template class Field<VData,ThreeD>; template class Field<VData,Sphere>;
template class Field<TData,Flat>; template class Field<TData,ThreeD>; template class Field<TData,Sphere>;
template class Field<QData,Flat>; template class Field<QData,ThreeD>; template class Field<QData,Sphere>;

namespace {
struct Dump {
    std::vector<double> pos; std::vector<float> size, w, wk, g;
    std::vector<long> n; std::vector<int> left, right, top;
};

// a complex field cell's w*z sum (C++11: no if constexpr)
template <int D, int C> struct WZ {
    static std::complex<double> get(const BaseCell<C>*) { return 0.; }
};
#define TC_WZ(DD) template <int C> struct WZ<DD, C> { \
    static std::complex<double> get(const BaseCell<C>* c) { \
        return static_cast<const Cell<DD,C>*>(c)->getWZ(); } };
TC_WZ(ZData) TC_WZ(VData) TC_WZ(GData) TC_WZ(TData) TC_WZ(QData)
#undef TC_WZ

template <int D, int C>
int walk(const BaseCell<C>* c, Dump& d) {
    int me = int(d.n.size());
    const Position<C>& p = c->getPos();
    d.pos.push_back(p.getX()); d.pos.push_back(p.getY());
    d.pos.push_back(C == Flat ? 0. : p.get(2));
    d.size.push_back(float(c->getSize())); d.w.push_back(float(c->getW()));
    float wk = 0, gr = 0, gi = 0;
    if (D == KData) wk = float(static_cast<const Cell<KData,C>*>(c)->getWK());
    if (D == ZData || D == VData || D == GData || D == TData || D == QData) {
        std::complex<double> z = WZ<D,C>::get(c);
        gr = float(z.real()); gi = float(z.imag());
    }
    d.wk.push_back(wk); d.g.push_back(gr); d.g.push_back(gi);
    d.n.push_back(c->getN()); d.left.push_back(-1); d.right.push_back(-1);
    if (c->getLeft()) {
        int l = walk<D,C>(c->getLeft(), d);
        int r = walk<D,C>(c->getRight(), d);
        d.left[me] = l; d.right[me] = r;
    }
    return me;
}

template <int D, int C>
Dump* build(const double* x, const double* y, const double* z, const double* d1,
            const double* d2, const double* w, const double* wpos, long n, double minsize,
            double maxsize, int sm, int brute, int mintop, int maxtop) {
    Field<D,C>* f = new Field<D,C>(x, y, z, d1, d2, w, wpos, n, minsize, maxsize,
                                   SplitMethod(sm), 0, brute != 0, mintop, maxtop);
    Dump* d = new Dump;
    const std::vector<const BaseCell<C>*>& cells = f->getCells();
    for (size_t i = 0; i < cells.size(); ++i) d->top.push_back(walk<D,C>(cells[i], *d));
    delete f;
    return d;
}
}  // namespace

extern "C" {
void* refdump_build(int D, int C, const double* x, const double* y, const double* z,
                    const double* d1, const double* d2, const double* w, const double* wpos,
                    long n, double minsize, double maxsize, int sm, int brute, int mintop,
                    int maxtop) {
#define RD(DD, CC) return build<DD, CC>(x, y, z, d1, d2, w, wpos, n, minsize, maxsize, sm, brute, mintop, maxtop)
    if (D == NData) { if (C == Flat) RD(NData, Flat); if (C == ThreeD) RD(NData, ThreeD); RD(NData, Sphere); }
    if (D == KData) { if (C == Flat) RD(KData, Flat); if (C == ThreeD) RD(KData, ThreeD); RD(KData, Sphere); }
    if (D == ZData) { if (C == Flat) RD(ZData, Flat); if (C == ThreeD) RD(ZData, ThreeD); RD(ZData, Sphere); }
    if (D == VData) { if (C == Flat) RD(VData, Flat); if (C == ThreeD) RD(VData, ThreeD); RD(VData, Sphere); }
    if (D == TData) { if (C == Flat) RD(TData, Flat); if (C == ThreeD) RD(TData, ThreeD); RD(TData, Sphere); }
    if (D == QData) { if (C == Flat) RD(QData, Flat); if (C == ThreeD) RD(QData, ThreeD); RD(QData, Sphere); }
    if (C == Flat) RD(GData, Flat); if (C == ThreeD) RD(GData, ThreeD); RD(GData, Sphere);
#undef RD
}
long refdump_nnodes(void* h) { return long(static_cast<Dump*>(h)->n.size()); }
long refdump_ntop(void* h) { return long(static_cast<Dump*>(h)->top.size()); }
void refdump_get(void* h, double* pos, float* size, float* w, float* wk, float* g, long* n,
                 int* left, int* right, int* top) {
    Dump* d = static_cast<Dump*>(h);
    for (size_t i = 0; i < d->n.size(); ++i) {
        pos[3*i] = d->pos[3*i]; pos[3*i+1] = d->pos[3*i+1]; pos[3*i+2] = d->pos[3*i+2];
        size[i] = d->size[i]; w[i] = d->w[i]; wk[i] = d->wk[i];
        g[2*i] = d->g[2*i]; g[2*i+1] = d->g[2*i+1]; n[i] = d->n[i];
        left[i] = d->left[i]; right[i] = d->right[i];
    }
    for (size_t i = 0; i < d->top.size(); ++i) top[i] = d->top[i];
}
void refdump_free(void* h) { delete static_cast<Dump*>(h); }
}
