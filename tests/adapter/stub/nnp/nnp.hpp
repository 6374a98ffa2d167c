// Stand-ins for the parts of ace's NNP tree the adapter reads (NNP, NNH, BasisR, BasisA, Cutoff, NN::ANN and
// the species Type; src/nnp/nnp.hpp, nnh.hpp, basis_radial.hpp, basis_angular.hpp, src/ml/nn.hpp), filled
// from a flat text dump by read_dump().  Test scaffolding.
#pragma once
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>
#include "Eigen/Dense"

class Cutoff {
public:
    int name_ = 0;
    double rc_ = 0.0;
    int name() const { return name_; }
    double rc() const { return rc_; }
};

class BasisR {
public:
    int name_ = 0;
    Cutoff cut_;
    std::vector<double> rs_, eta_;
    int name() const { return name_; }
    const Cutoff& cutoff() const { return cut_; }
    int size() const { return (int)rs_.size(); }
    const std::vector<double>& rs() const { return rs_; }
    const std::vector<double>& eta() const { return eta_; }
};

class BasisA {
public:
    int name_ = 0;
    Cutoff cut_;
    std::vector<double> eta_, zeta_;
    std::vector<int> lambda_;
    int name() const { return name_; }
    const Cutoff& cutoff() const { return cut_; }
    int size() const { return (int)eta_.size(); }
    const std::vector<double>& eta() const { return eta_; }
    const std::vector<double>& zeta() const { return zeta_; }
    const std::vector<int>& lambda() const { return lambda_; }
};

namespace NN {
class ANN {
public:
    std::vector<int> widths_;            // [nInput, h1, ..., 1]
    int neuron_ = 0;
    double c_ = 1.0;
    Eigen::VectorXd inpw_, inpb_, outw_, outb_;
    std::vector<double> params_;         // [biases][weights column-major], the order of operator>>
    int nlayer() const { return (int)widths_.size() - 1; }
    int nInp() const { return widths_[0]; }
    int nNodes(int l) const { return widths_[l + 1]; }
    int neuron() const { return neuron_; }
    double c() const { return c_; }
    const Eigen::VectorXd& inpw() const { return inpw_; }
    const Eigen::VectorXd& inpb() const { return inpb_; }
    const Eigen::VectorXd& outw() const { return outw_; }
    const Eigen::VectorXd& outb() const { return outb_; }
    const ANN& operator>>(Eigen::VectorXd& v) const {
        v.d = params_;
        return *this;
    }
};
}  // namespace NN

class Type {
public:
    double energy_ = 0.0;
    double energy() const { return energy_; }
};

class NNH {
public:
    int ntypes_ = 0;
    std::vector<BasisR> basisR_;
    std::vector<BasisA> basisA_;         // TensorTri order: index K (K + 1) / 2 + J for J <= K
    NN::ANN nn_;
    Type type_;
    const BasisR& basisR(int J) const { return basisR_[J]; }
    const BasisA& basisA(int J, int K) const { return basisA_[K * (K + 1) / 2 + J]; }
    const NN::ANN& nn() const { return nn_; }
    const Type& type() const { return type_; }
};

class NNP {
public:
    int ntypes_ = 0;
    double rc_ = 0.0;
    std::vector<NNH> nnh_;
    int ntypes() const { return ntypes_; }
    double rc() const { return rc_; }
    const NNH& nnh(int I) const { return nnh_[I]; }
};

// Dump format (whitespace separated): ntypes rc; per species I: per J radial "form cut rc n rs[n] eta[n]";
// per TensorTri block angular "form cut rc n eta[n] zeta[n] lambda[n]"; network "nlayer widths[nlayer+1] act c
// inpw[K] inpb[K] outw outb E0 nparams params[nparams]".
inline NNP read_dump(const std::string& file) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open " + file);
    NNP p;
    in >> p.ntypes_ >> p.rc_;
    const int nt = p.ntypes_, nab = nt * (nt + 1) / 2;
    p.nnh_.resize(nt);
    for (int I = 0; I < nt; ++I) {
        NNH& h = p.nnh_[I];
        h.ntypes_ = nt;
        h.basisR_.resize(nt);
        for (BasisR& b : h.basisR_) {
            int n;
            in >> b.name_ >> b.cut_.name_ >> b.cut_.rc_ >> n;
            b.rs_.resize(n);
            b.eta_.resize(n);
            for (double& x : b.rs_) in >> x;
            for (double& x : b.eta_) in >> x;
        }
        h.basisA_.resize(nab);
        for (BasisA& b : h.basisA_) {
            int n;
            in >> b.name_ >> b.cut_.name_ >> b.cut_.rc_ >> n;
            b.eta_.resize(n);
            b.zeta_.resize(n);
            b.lambda_.resize(n);
            for (double& x : b.eta_) in >> x;
            for (double& x : b.zeta_) in >> x;
            for (int& x : b.lambda_) in >> x;
        }
        NN::ANN& a = h.nn_;
        int nl;
        in >> nl;
        a.widths_.resize(nl + 1);
        for (int& w : a.widths_) in >> w;
        in >> a.neuron_ >> a.c_;
        const int K = a.widths_[0];
        a.inpw_.d.resize(K);
        a.inpb_.d.resize(K);
        for (double& x : a.inpw_.d) in >> x;
        for (double& x : a.inpb_.d) in >> x;
        a.outw_.d.resize(1);
        a.outb_.d.resize(1);
        in >> a.outw_.d[0] >> a.outb_.d[0] >> h.type_.energy_;
        int np;
        in >> np;
        a.params_.resize(np);
        for (double& x : a.params_) in >> x;
    }
    if (!in) throw std::runtime_error("truncated model dump " + file);
    return p;
}
