// cell_grid.hpp -- the walk over grid points shared by the host-side
// initialisers (Psi, Vext).
//
// A point's position is its cell centre  -XMAX + i DX + DX/2  per axis (the
// reference's getX/getY/getZ, src/psi.cu:134-150, src/vext.cu:182-199).  The
// centres are tabulated once per axis with that one expression, so every
// initialiser sees the same bits the per-point formula gives, and the walk
// visits points in storage order (x fastest, index z*NX*NY + y*NX + x) so
// that sequential state (the noise stream, running sums) advances as in the
// reference's index loops.
#ifndef FGPE_HOST_CELL_GRID_HPP
#define FGPE_HOST_CELL_GRID_HPP

#include <vector>

#include "../hpp/params.hpp"

namespace fgpe_host {

inline double cell_centre(const ParamsBase& p, int axis, int i)
{
    return -p.XMAX[axis] + i * p.DX[axis] + 0.5 * p.DX[axis];
}

class CellGrid {
public:
    explicit CellGrid(const ParamsBase& p)
    {
        for (int a = 0; a < 3; ++a) {
            c_[a].resize(p.NX[a]);
            for (int i = 0; i < p.NX[a]; ++i) c_[a][i] = cell_centre(p, a, i);
        }
    }

    // f(i, x, y, z) for every point, i the storage index
    template <class F> void walk(F&& f) const
    {
        long long i = 0;
        for (double z : c_[2])
            for (double y : c_[1])
                for (double x : c_[0]) f(i++, x, y, z);
    }

private:
    std::vector<double> c_[3];
};

// dx dy dz of the grid
inline double cell_volume(const ParamsBase& p) { return p.DX[0] * p.DX[1] * p.DX[2]; }

}  // namespace fgpe_host

#endif
