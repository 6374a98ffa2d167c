// CalcNNP: the MD Calculator of the integration guide (sim/calc.hpp interface), compiled here as it would be
// inside ace (against the stand-in types of tests/adapter/stub).
#pragma once
#include <stdexcept>
#include <vector>
#include "nnp_device.hpp"
#include "sim/calc.hpp"
class CalcNNP: public Calculator {
    NNPDevice* dev_;                              // the adapter above (owns the ace_ctx)
    mutable std::vector<double> pos_, F_;
public:
    CalcNNP(NNPDevice* dev, double rc): Calculator(Name::NNP, rc), dev_(dev) {}
    void init(const Structure& struc) {           // once: species, cell, initial positions
        const int n = struc.nAtoms();
        pos_.resize(3 * n); F_.resize(3 * n);
        std::vector<int> type(n);
        std::vector<double> R(9);
        for (int i = 0; i < n; ++i) { type[i] = struc.type(i); for (int d = 0; d < 3; ++d) pos_[3 * i + d] = struc.posn(i)[d]; }
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) R[3 * c + r] = struc.R()(r, c);
        if (ace_upload(dev_->ctx(), 1, &n, pos_.data(), type.data(), R.data()))
            throw std::runtime_error(ace_last_error(dev_->ctx()));
    }
    double compute(Structure& struc) const {      // positions in, E and F out (forces added)
        const int n = struc.nAtoms();
        for (int i = 0; i < n; ++i) for (int d = 0; d < 3; ++d) pos_[3 * i + d] = struc.posn(i)[d];
        double E = 0;
        if (ace_update_positions(dev_->ctx(), pos_.data()) ||
            ace_compute(dev_->ctx(), ACE_PREC_FP32, 0.0, &E, F_.data(), nullptr))
            throw std::runtime_error(ace_last_error(dev_->ctx()));
        for (int i = 0; i < n; ++i) struc.force(i) += Eigen::Vector3d(F_[3 * i], F_[3 * i + 1], F_[3 * i + 2]);
        return E;
    }
    double compute(Structure& struc, const NeighborList&) const { return compute(struc); }
    double energy(Structure& struc) const { return compute(struc); }
};
