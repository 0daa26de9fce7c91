// Gate fusion (src/fusion.rs).
//
// Phase 1 (:14-134): absorb every single-qubit gate into its next (else
// previous) neighbour on the same wire, then fuse adjacent two-qubit gates
// on identical qubit pairs.  Phase 2 (:168-345): agglomerative fusion for
// candidate sizes 3..size_max, each size repeated until a pass fuses nothing;
// a cluster is accepted when sum(old costs) / (new cost + 1e-10) - 1 >=
// benefit_margin, otherwise the original gates are restored.
#pragma once

#include "circuit.hpp"
#include "cost_model.hpp"

namespace cast {

void apply_size_two_fusion(CircuitGraph &cg);
size_t apply_gate_fusion(CircuitGraph &cg, const FusionConfig &cfg, size_t cdd_size,
                         FusionLog *log);
void optimize(CircuitGraph &cg, const FusionConfig &cfg);
FusionLog optimize_with_log(CircuitGraph &cg, const FusionConfig &cfg);

} // namespace cast
