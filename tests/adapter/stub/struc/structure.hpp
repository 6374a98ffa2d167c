// Stand-in for ace's Structure (src/struc/structure.hpp): the accessors the adapter uses.  Test scaffolding.
#pragma once
#include <vector>
#include "Eigen/Dense"

class Structure {
    std::vector<Eigen::Vector3d> posn_, force_;
    std::vector<int> type_;
    Eigen::Matrix3d R_;
    double pe_ = 0.0;

public:
    void resize(int n) {
        posn_.assign(n, Eigen::Vector3d());
        force_.assign(n, Eigen::Vector3d());
        type_.assign(n, 0);
    }
    int nAtoms() const { return (int)posn_.size(); }
    Eigen::Vector3d& posn(int i) { return posn_[i]; }
    const Eigen::Vector3d& posn(int i) const { return posn_[i]; }
    Eigen::Vector3d& force(int i) { return force_[i]; }
    int& type(int i) { return type_[i]; }
    int type(int i) const { return type_[i]; }
    Eigen::Matrix3d& R() { return R_; }
    const Eigen::Matrix3d& R() const { return R_; }
    double& pe() { return pe_; }
};

class NeighborList {};
