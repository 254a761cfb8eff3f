// Query: a plan tree built from a libconfig plan by type string, driven
// through the lifecycle as Query does in the reference (query.h:40-135,
// query.cpp:37-247).
#pragma once

#include "operators.h"

class Query {
  public:
    void threadInit() {
        ThreadInitVisitor v(0);
        tree->accept(&v);
    }
    Operator::ResultCode scanStart() {
        Schema empty;
        return tree->scanStart(0, nullptr, empty);
    }
    Operator::GetNextResultT getNext() { return tree->getNext(0); }
    Operator::ResultCode scanStop() { return tree->scanStop(0); }
    void threadClose() {
        ThreadCloseVisitor v(0);
        tree->accept(&v);
    }
    void destroynofree() {
        RecursiveDestroyVisitor v;
        tree->accept(&v);
    }
    void destroy() {
        RecursiveDestroyVisitor d;
        tree->accept(&d);
        RecursiveFreeVisitor f;
        tree->accept(&f);
        tree = nullptr;
    }
    Schema& getOutSchema() { return tree->getOutSchema(); }
    // Builds the tree under `treeroot` and calls init bottom-up.
    void create(libconfig::Config& cfg);

    Operator* tree = nullptr;
};
