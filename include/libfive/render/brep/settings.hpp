// Meshing settings: the members of libfive::BRepSettings
// (libfive/include/libfive/render/brep/settings.hpp:24-95) with the same
// names, types and defaults.  This library implements the DUAL_CONTOURING
// algorithm; min_feature, max_err and cancel are honoured, workers is
// accepted (the GPU has its own parallelism), and the progress and
// free-thread handlers and the volume tree are not used.
#pragma once

#include <atomic>

namespace libfive {

class ProgressHandler;
class FreeThreadHandler;
class VolTree;

enum BRepAlgorithm {
    DUAL_CONTOURING,
    ISO_SIMPLEX,
    HYBRID,
};

struct BRepSettings {
    BRepSettings() { reset(); }
    void reset() {
        min_feature = 0.1;
        max_err = 1e-8;
        workers = 8;
        alg = DUAL_CONTOURING;
        free_thread_handler = nullptr;
        progress_handler = nullptr;
        cancel.store(false);
        vol = nullptr;
    }
    double min_feature;      // subdivide until a cell edge is below this
    double max_err;          // QEF error bound for collapsing cells
    unsigned workers;
    BRepAlgorithm alg;
    FreeThreadHandler* free_thread_handler;
    ProgressHandler* progress_handler;
    const VolTree* vol;
    mutable std::atomic_bool cancel;   // set from any thread to stop a render
};

}   // namespace libfive
