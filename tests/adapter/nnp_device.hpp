// NNPDevice: the C++ adapter of the integration guide (drop-in for NNP::compute callers), compiled here as it
// would be inside ace (against the stand-in types of tests/adapter/stub).
#pragma once
#include <stdexcept>
#include <vector>
#include "ace_hip.h"
#include "nnp/nnp.hpp"
#include "struc/structure.hpp"

class NNPDevice {
    ace_ctx* ctx_ = nullptr;
    std::vector<int> rf_, rc_, rn_, ro_, af_, ac_, an_, ao_, al_, nl_, nw_, na_, io_, po_;
    std::vector<double> rrc_, rrs_, reta_, arc_, aeta_, azeta_, nc_, iw_, ib_, ow_, ob_, e0_, p_;
public:
    ace_ctx* ctx() const { return ctx_; }
    explicit NNPDevice(const NNP& nnp) {
        ace_model_desc d{};
        d.ntypes = nnp.ntypes();
        d.rc = nnp.rc();
        for (int I = 0; I < nnp.ntypes(); ++I) {
            const NNH& h = nnp.nnh(I);
            for (int J = 0; J < nnp.ntypes(); ++J) {          // BasisR blocks
                const BasisR& b = h.basisR(J);
                rf_.push_back(b.name()); rc_.push_back(b.cutoff().name()); rrc_.push_back(b.cutoff().rc());
                rn_.push_back(b.size()); ro_.push_back(rrs_.size());
                rrs_.insert(rrs_.end(), b.rs().begin(), b.rs().end());
                reta_.insert(reta_.end(), b.eta().begin(), b.eta().end());
            }
            for (int K = 0; K < nnp.ntypes(); ++K)            // BasisA blocks, TensorTri order
                for (int J = 0; J <= K; ++J) {
                    const BasisA& b = h.basisA(J, K);
                    af_.push_back(b.name()); ac_.push_back(b.cutoff().name()); arc_.push_back(b.cutoff().rc());
                    an_.push_back(b.size()); ao_.push_back(aeta_.size());
                    aeta_.insert(aeta_.end(), b.eta().begin(), b.eta().end());
                    azeta_.insert(azeta_.end(), b.zeta().begin(), b.zeta().end());
                    al_.insert(al_.end(), b.lambda().begin(), b.lambda().end());
                }
            const NN::ANN& nn = h.nn();                       // network, inputs scaling, parameters
            nl_.push_back(nn.nlayer());
            std::vector<int> w(ACE_MAX_LAYERS + 1, 0);
            w[0] = nn.nInp();
            for (int l = 0; l < nn.nlayer(); ++l) w[l + 1] = nn.nNodes(l);
            nw_.insert(nw_.end(), w.begin(), w.end());
            na_.push_back(nn.neuron()); nc_.push_back(nn.c());
            io_.push_back(iw_.size());
            iw_.insert(iw_.end(), nn.inpw().data(), nn.inpw().data() + nn.nInp());
            ib_.insert(ib_.end(), nn.inpb().data(), nn.inpb().data() + nn.nInp());
            ow_.push_back(nn.outw()[0]); ob_.push_back(nn.outb()[0]); e0_.push_back(h.type().energy());
            Eigen::VectorXd v; nn >> v;                       // [biases][weights col-major]
            po_.push_back(p_.size()); p_.insert(p_.end(), v.data(), v.data() + v.size());
        }
        d.r_form = rf_.data(); d.r_cut = rc_.data(); d.r_rc = rrc_.data(); d.r_n = rn_.data(); d.r_off = ro_.data();
        d.r_rs = rrs_.data(); d.r_eta = reta_.data();
        d.a_form = af_.data(); d.a_cut = ac_.data(); d.a_rc = arc_.data(); d.a_n = an_.data(); d.a_off = ao_.data();
        d.a_eta = aeta_.data(); d.a_zeta = azeta_.data(); d.a_lambda = al_.data();
        d.nn_nlayer = nl_.data(); d.nn_width = nw_.data(); d.nn_act = na_.data(); d.nn_c = nc_.data();
        d.in_off = io_.data(); d.inpw = iw_.data(); d.inpb = ib_.data(); d.outw = ow_.data(); d.outb = ob_.data();
        d.E0 = e0_.data(); d.par_off = po_.data(); d.params = p_.data();
        if (ace_create(&ctx_, &d)) throw std::invalid_argument(ace_last_error(ctx_));
    }
    ~NNPDevice() { ace_destroy(ctx_); }

    // NNP::compute(Structure&): neighbour list, G, energy and forces written back into struc
    double compute(Structure& struc) {
        const int n = struc.nAtoms();
        std::vector<double> pos(3 * n), F(3 * n), R(9);
        std::vector<int> type(n);
        for (int i = 0; i < n; ++i) {
            for (int d = 0; d < 3; ++d) pos[3 * i + d] = struc.posn(i)[d];
            type[i] = struc.type(i);
        }
        for (int c = 0; c < 3; ++c) for (int r = 0; r < 3; ++r) R[3 * c + r] = struc.R()(r, c);
        double E = 0;
        if (ace_upload(ctx_, 1, &n, pos.data(), type.data(), R.data()) ||
            ace_compute(ctx_, ACE_PREC_FP32, 0.0, &E, F.data(), nullptr))
            throw std::runtime_error(ace_last_error(ctx_));
        for (int i = 0; i < n; ++i) struc.force(i) << F[3 * i], F[3 * i + 1], F[3 * i + 2];
        struc.pe() = E;
        return E;
    }
};
