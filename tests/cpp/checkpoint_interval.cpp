// Host-only check of physgrad::mpm::differentiable::checkpointInterval (no
// device calls): prints the interval for each case given on the command line
// as "strategy steps budget_gb max_checkpoints use_checkpointing every ckpt_bytes step_bytes".
#include <cstdio>
#include <cstdlib>

#include "physgrad/differentiable_mpm.h"

using namespace physgrad::mpm::differentiable;

int main(int argc, char** argv) {
    for (int i = 1; i + 7 < argc; i += 8) {
        DifferentiableMPMConfig<float> c;
        c.checkpoint_strategy = static_cast<CheckpointStrategy>(std::atoi(argv[i]));
        const int steps = std::atoi(argv[i + 1]);
        c.memory_budget_gb = static_cast<float>(std::atof(argv[i + 2]));
        c.max_checkpoints = static_cast<size_t>(std::atol(argv[i + 3]));
        c.use_checkpointing = std::atoi(argv[i + 4]) != 0;
        c.checkpoint_every = std::atoi(argv[i + 5]);
        std::printf("%d\n", checkpointInterval(c, steps, std::atof(argv[i + 6]), std::atof(argv[i + 7])));
    }
    return 0;
}
