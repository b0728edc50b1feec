// StateVars.h -- {t, q, aux} of the FINESS app interface (lib/2d/StateVars.h:
// 9-53, lib/1d/StateVars.h), written for this library.  aux holds
// max(1, maux) fields.  The 1D form is selected with NDIMS=1.
#ifndef FINESS_COMPAT_STATEVARS_H
#define FINESS_COMPAT_STATEVARS_H

#include <algorithm>
#include "tensors.h"

class StateVars {
public:
#if NDIMS == 1
    typedef dTensorBC2 Field;
    StateVars(double t, int mx, int meqn, int maux, int mbc)
        : t_(t), q_(mx, meqn, mbc), aux_(mx, std::max(1, maux), mbc) {}
#else
    typedef dTensorBC3 Field;
    StateVars(double t, int mx, int my, int meqn, int maux, int mbc)
        : t_(t), q_(mx, my, meqn, mbc), aux_(mx, my, std::max(1, maux), mbc) {}
#endif
    double get_t() const { return t_; }
    void set_t(double t) { t_ = t; }
    const Field& const_ref_q() const { return q_; }
    Field& ref_q() { return q_; }
    const Field& const_ref_aux() const { return aux_; }
    Field& ref_aux() { return aux_; }
    void copyfrom(const StateVars& o)
    {
        q_.copyfrom(o.q_);
        aux_.copyfrom(o.aux_);
        t_ = o.t_;
    }
private:
    double t_;
    Field q_, aux_;
};

#endif
