// CommonParams.h -- lattice/run parameters handed to a model, field for
// field as the reference's struct (src/utils/CommonParams.h:5-23).  One GPU
// holds the whole lattice, so the local sizes equal the global ones, rank is
// 0 of np = 1 and both x neighbours are rank 0.
#pragma once

namespace mesobasic {

struct CommonParams {
    int nx, ny, nz;     // nodes in the local domain
    int NX, NY, NZ;     // nodes in the global domain
    int rank;
    int np;
    int nbrL;
    int nbrR;
    int xOff;
    int iskip;
    int jskip;
    int kskip;
    int nstep;
    int numOutputs;
    double dx, dy, dz;
    double LX, LY, LZ;
    double dt;
};

}  // namespace mesobasic
