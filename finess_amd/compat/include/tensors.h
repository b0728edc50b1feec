// tensors.h -- the tensor containers of the FINESS app interface, written for
// this library (interface of lib/tensors.h / lib/tensors1d.h): 1-based
// indices, row-major storage with the last index fastest, and mbc ghost
// layers on the first `ndims` indices of the dTensorBC* classes, so that a
// dTensorBC3 q(mx,my,meqn,mbc) stores q(i,j,m) at
// ((i+mbc-1)*(my+2mbc) + (j+mbc-1))*meqn + (m-1) (lib/tensors.cpp:650-667).
// Storage is value-initialised (zero), unlike the reference's new double[].
#ifndef FINESS_COMPAT_TENSORS_H
#define FINESS_COMPAT_TENSORS_H

#include <cassert>
#include <cstring>
#include <vector>

#ifndef NDIMS
#define NDIMS 2
#endif

class dTensorBase {
public:
    explicit dTensorBase(int n = 0) : v_(n > 0 ? n : 0, 0.0) {}
    int numel() const { return (int)v_.size(); }
    const double& vget(int k) const { return v_[k]; }
    void vset(int k, double x) { v_[k] = x; }
    void setall(double x) { std::fill(v_.begin(), v_.end(), x); }
    double* data() { return v_.data(); }
    const double* data() const { return v_.data(); }
protected:
    void copy_storage(const dTensorBase& o) { assert(o.v_.size() == v_.size()); v_ = o.v_; }
    std::vector<double> v_;
};

// general N-index tensor: sizes s[], first indices b[]
template <int N>
class dTensorNd : public dTensorBase {
public:
    dTensorNd(const int* s, const int* b) : dTensorBase(prod(s))
    {
        for (int k = 0; k < N; k++) { s_[k] = s[k]; b_[k] = b[k]; }
    }
    int idx(const int* n) const
    {
        int k = n[0] - b_[0];
        for (int d = 1; d < N; d++) k = k * s_[d] + (n[d] - b_[d]);
        return k;
    }
protected:
    static int prod(const int* s) { int p = 1; for (int k = 0; k < N; k++) p *= s[k]; return p; }
    int s_[N], b_[N];
};

class dTensor1 : public dTensorNd<1> {
public:
    explicit dTensor1(int s1) : dTensorNd<1>(sz(s1), one()) {}
    const double& get(int i) const { return v_[i - 1]; }
    void set(int i, double x) { v_[i - 1] = x; }
    int getsize() const { return s_[0]; }
    int getsize(int) const { return s_[0]; }
    void copyfrom(const dTensor1& o) { copy_storage(o); }
private:
    static const int* sz(int a) { static thread_local int s[1]; s[0] = a; return s; }
    static const int* one() { static const int b[1] = {1}; return b; }
};

#define FINESS_TENSOR_ACCESSORS(N)                                                       \
    int getsize(int d) const { return (d >= 1 && d <= N) ? S_[d - 1] : 1; }

class dTensor2 : public dTensorNd<2> {
public:
    dTensor2(int s1, int s2) : dTensorNd<2>(mk(s1, s2), ones()) { S_[0] = s1; S_[1] = s2; }
    const double& get(int i, int j) const { return v_[(i - 1) * s_[1] + (j - 1)]; }
    void set(int i, int j, double x) { v_[(i - 1) * s_[1] + (j - 1)] = x; }
    FINESS_TENSOR_ACCESSORS(2)
    void copyfrom(const dTensor2& o) { copy_storage(o); }
private:
    static const int* mk(int a, int b) { static thread_local int s[2]; s[0] = a; s[1] = b; return s; }
    static const int* ones() { static const int b[2] = {1, 1}; return b; }
    int S_[2];
};

class dTensor3 : public dTensorNd<3> {
public:
    dTensor3(int s1, int s2, int s3) : dTensorNd<3>(mk(s1, s2, s3), ones()) { S_[0] = s1; S_[1] = s2; S_[2] = s3; }
    const double& get(int i, int j, int k) const { return v_[((i - 1) * s_[1] + (j - 1)) * s_[2] + (k - 1)]; }
    void set(int i, int j, int k, double x) { v_[((i - 1) * s_[1] + (j - 1)) * s_[2] + (k - 1)] = x; }
    FINESS_TENSOR_ACCESSORS(3)
    void copyfrom(const dTensor3& o) { copy_storage(o); }
private:
    static const int* mk(int a, int b, int c) { static thread_local int s[3]; s[0] = a; s[1] = b; s[2] = c; return s; }
    static const int* ones() { static const int b[3] = {1, 1, 1}; return b; }
    int S_[3];
};

class dTensor4 : public dTensorNd<4> {
public:
    dTensor4(int s1, int s2, int s3, int s4) : dTensorNd<4>(mk(s1, s2, s3, s4), ones())
    { S_[0] = s1; S_[1] = s2; S_[2] = s3; S_[3] = s4; }
    const double& get(int a, int b, int c, int d) const { int n[4] = {a, b, c, d}; return v_[idx(n)]; }
    void set(int a, int b, int c, int d, double x) { int n[4] = {a, b, c, d}; v_[idx(n)] = x; }
    FINESS_TENSOR_ACCESSORS(4)
    void copyfrom(const dTensor4& o) { copy_storage(o); }
private:
    static const int* mk(int a, int b, int c, int d) { static thread_local int s[4]; s[0] = a; s[1] = b; s[2] = c; s[3] = d; return s; }
    static const int* ones() { static const int b[4] = {1, 1, 1, 1}; return b; }
    int S_[4];
};

class dTensor5 : public dTensorNd<5> {
public:
    dTensor5(int s1, int s2, int s3, int s4, int s5) : dTensorNd<5>(mk(s1, s2, s3, s4, s5), ones())
    { S_[0] = s1; S_[1] = s2; S_[2] = s3; S_[3] = s4; S_[4] = s5; }
    const double& get(int a, int b, int c, int d, int e) const { int n[5] = {a, b, c, d, e}; return v_[idx(n)]; }
    void set(int a, int b, int c, int d, int e, double x) { int n[5] = {a, b, c, d, e}; v_[idx(n)] = x; }
    FINESS_TENSOR_ACCESSORS(5)
    void copyfrom(const dTensor5& o) { copy_storage(o); }
private:
    static const int* mk(int a, int b, int c, int d, int e)
    { static thread_local int s[5]; s[0] = a; s[1] = b; s[2] = c; s[3] = d; s[4] = e; return s; }
    static const int* ones() { static const int b[5] = {1, 1, 1, 1, 1}; return b; }
    int S_[5];
};

// dTensorBC1(s1, mbc): ghosts on index 1.
class dTensorBC1 : public dTensorBase {
public:
    dTensorBC1(int s1, int mbc, int = 1) : dTensorBase(s1 + 2 * mbc), S1(s1), mbc_(mbc) {}
    const double& get(int i) const { return v_[i + mbc_ - 1]; }
    void set(int i, double x) { v_[i + mbc_ - 1] = x; }
    int getsize(int = 1) const { return S1; }
    int getmbc() const { return mbc_; }
    void copyfrom(const dTensorBC1& o) { copy_storage(o); }
private:
    int S1, mbc_;
};

// dTensorBC2(s1, s2, mbc, ndims=NDIMS): ghosts on the first ndims indices.
class dTensorBC2 : public dTensorBase {
public:
    dTensorBC2(int s1, int s2, int mbc, int ndims = NDIMS)
        : dTensorBase((s1 + 2 * mbc) * (s2 + (ndims >= 2 ? 2 * mbc : 0))), S1(s1), S2(s2), mbc_(mbc),
          g2(ndims >= 2 ? mbc : 0) {}
    const double& get(int i, int j) const { return v_[(i + mbc_ - 1) * (S2 + 2 * g2) + (j + g2 - 1)]; }
    void set(int i, int j, double x) { v_[(i + mbc_ - 1) * (S2 + 2 * g2) + (j + g2 - 1)] = x; }
    int getsize(int d) const { return d == 1 ? S1 : (d == 2 ? S2 : 1); }
    int getmbc() const { return mbc_; }
    void copyfrom(const dTensorBC2& o) { copy_storage(o); }
private:
    int S1, S2, mbc_, g2;
};

// dTensorBC3(s1, s2, s3, mbc, ndims=NDIMS)
class dTensorBC3 : public dTensorBase {
public:
    dTensorBC3(int s1, int s2, int s3, int mbc, int ndims = NDIMS)
        : dTensorBase((s1 + 2 * mbc) * (s2 + (ndims >= 2 ? 2 * mbc : 0)) * (s3 + (ndims >= 3 ? 2 * mbc : 0))),
          S1(s1), S2(s2), S3(s3), mbc_(mbc), g2(ndims >= 2 ? mbc : 0), g3(ndims >= 3 ? mbc : 0) {}
    int ix(int i, int j, int k) const
    { return ((i + mbc_ - 1) * (S2 + 2 * g2) + (j + g2 - 1)) * (S3 + 2 * g3) + (k + g3 - 1); }
    const double& get(int i, int j, int k) const { return v_[ix(i, j, k)]; }
    void set(int i, int j, int k, double x) { v_[ix(i, j, k)] = x; }
    int getsize(int d) const { return d == 1 ? S1 : (d == 2 ? S2 : (d == 3 ? S3 : 1)); }
    int getmbc() const { return mbc_; }
    void copyfrom(const dTensorBC3& o) { copy_storage(o); }
private:
    int S1, S2, S3, mbc_, g2, g3;
};

// dTensorBC4(s1, s2, s3, s4, mbc, ndims=NDIMS)
class dTensorBC4 : public dTensorBase {
public:
    dTensorBC4(int s1, int s2, int s3, int s4, int mbc, int ndims = NDIMS)
        : dTensorBase((s1 + 2 * mbc) * (s2 + (ndims >= 2 ? 2 * mbc : 0)) * s3 * s4), S1(s1), S2(s2), S3(s3), S4(s4),
          mbc_(mbc), g2(ndims >= 2 ? mbc : 0) {}
    int ix(int i, int j, int k, int l) const
    { return (((i + mbc_ - 1) * (S2 + 2 * g2) + (j + g2 - 1)) * S3 + (k - 1)) * S4 + (l - 1); }
    const double& get(int i, int j, int k, int l) const { return v_[ix(i, j, k, l)]; }
    void set(int i, int j, int k, int l, double x) { v_[ix(i, j, k, l)] = x; }
    int getsize(int d) const { return d == 1 ? S1 : d == 2 ? S2 : d == 3 ? S3 : d == 4 ? S4 : 1; }
    int getmbc() const { return mbc_; }
    void copyfrom(const dTensorBC4& o) { copy_storage(o); }
private:
    int S1, S2, S3, S4, mbc_, g2;
};

#endif
