// Stand-in for ace's Calculator base (src/sim/calc.hpp:17-108) with NNP added to its names, as a maintainer
// adding the device calculator would.  Test scaffolding.
#pragma once
#include "struc/structure.hpp"

class Calculator {
public:
    class Name {
    public:
        enum Type { LJ_CUT, COUL_CUT, NNP, NONE };
        Name() : t_(NONE) {}
        Name(Type t) : t_(t) {}
        operator Type() const { return t_; }

    private:
        Type t_;
    };

protected:
    Name name_;
    double rc_ = 0.0, rc2_ = 0.0;

public:
    Calculator(Name name, double rc) : name_(name), rc_(rc), rc2_(rc * rc) {}
    virtual ~Calculator() {}
    const Name& name() const { return name_; }
    double rc() const { return rc_; }
    virtual void init(const Structure&) {}
    virtual double energy(Structure&) const { return 0.0; }
    virtual double compute(Structure&) const { return 0.0; }
    virtual double compute(Structure&, const NeighborList&) const { return 0.0; }
};
