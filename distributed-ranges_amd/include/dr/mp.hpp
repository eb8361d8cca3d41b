// dr/mp.hpp -- the multi-process model's sparse gemv interface on MI355X:
// runtime (ranks, communicator), broadcasted right-hand sides, distributed
// CSR matrix with csr_row / csr_eq distributions, and dr::mp::gemv.
#pragma once

#include "dr/concepts.hpp"
#include "dr/detail/generate_csr.hpp"
#include "dr/detail/index.hpp"
#include "dr/detail/matrix_io.hpp"
#include "dr/mp/broadcasted.hpp"
#include "dr/mp/gemv.hpp"
#include "dr/mp/runtime.hpp"
#include "dr/mp/sparse.hpp"
#include "dr/views/csr_matrix_view.hpp"
