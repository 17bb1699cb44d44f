"""Layer stacks shared by the oracle and GPU tests of the general
FeedForward network: (n, layers), layers as tuples (kind, a, b, c, d,
use_bias, param) in addLayer order (the numbering of yq_ff_layer)."""
import oracle as orc

C1, FC, ACT, MEAN = orc.CONV1D, orc.FC, orc.ACT, orc.MEAN

# (n, layers): tuples (kind, a, b, c, d, use_bias, param)
STACKS = {
    # Examples/AmplitudePhase.cpp:42-55 at N = 12
    "amplitude_phase_example": (12, [
        (C1, 1, 12, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
        (C1, 12, 10, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
        (C1, 10, 8, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
        (C1, 8, 6, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
        (C1, 6, 4, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
        (C1, 4, 2, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
        (FC, 24, 1, 0, 0, 0, 0), (ACT, orc.A_SOFTSIGN, 0, 0, 0, 0, 0)]),
    # channels, stride 2, biases everywhere
    "stride_bias": (10, [
        (C1, 1, 4, 3, 1, 1, 0), (ACT, orc.A_TANH, 0, 0, 0, 0, 2.0),
        (C1, 4, 3, 4, 2, 1, 0), (ACT, orc.A_RELU, 0, 0, 0, 0, 0),
        (FC, 15, 5, 0, 0, 1, 0), (ACT, orc.A_LEAKYRELU, 0, 0, 0, 0, 0.1),
        (FC, 5, 1, 0, 0, 1, 0), (ACT, orc.A_IDENTITY, 0, 0, 0, 0, 0)]),
    # the other activations and Mean over both axes
    "activations_mean": (9, [
        (C1, 1, 3, 3, 1, 0, 0), (ACT, orc.A_LNCOSH, 0, 0, 0, 0, 0),
        (ACT, orc.A_WEAKTANH, 0, 0, 0, 0, 0), (MEAN, 9, 3, 1, 0, 0, 0),
        (FC, 9, 4, 0, 0, 1, 0), (ACT, orc.A_HARDTANH, 0, 0, 0, 0, 0),
        (FC, 4, 2, 0, 0, 1, 0), (ACT, orc.A_SOFTSHRINK, 0, 0, 0, 0, 0),
        (ACT, orc.A_COS, 0, 0, 0, 0, 1.5), (MEAN, 2, 1, 0, 0, 0, 0)]),
    # a kernel wider than half the chain and a stride that does not divide it
    "wide_kernel_odd_stride": (4, [
        (C1, 1, 2, 5, 1, 0, 0), (ACT, orc.A_TANH, 0, 0, 0, 0, 1.0),
        (C1, 2, 2, 3, 3, 1, 0), (ACT, orc.A_LNCOSH, 0, 0, 0, 0, 0),
        (MEAN, 1, 2, 1, 0, 0, 0)]),
    # the C4 network as a stack
    "c4": (8, [
        (C1, 1, 3, 3, 1, 0, 0), (ACT, orc.A_LNCOSH, 0, 0, 0, 0, 0),
        (FC, 24, 6, 0, 0, 0, 0), (ACT, orc.A_TANH, 0, 0, 0, 0, 1.0),
        (MEAN, 6, 1, 0, 0, 0, 0)]),
}

# Sigmoid as the reference writes it (activations.hpp:211-230): the oracle's
# restatement of that backward is not a derivative, so this stack is compared
# with the oracle on the GPU only
SIGMOID_STACK = (6, [
    (FC, 6, 3, 0, 0, 1, 0), (ACT, orc.A_SIGMOID, 0, 0, 0, 0, 0),
    (FC, 3, 1, 0, 0, 0, 0)])

# Examples/AmplitudePhase.cpp:42-55 at its own N = 40 (Conv1D 1 -> 12 -> 10 ->
# 8 -> 6 -> 4 -> 2 with LeakyHardTanh(0.01), FullyConnected(2N, 1), SoftSign)
AP_EXAMPLE_40 = (40, [
    (C1, 1, 12, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
    (C1, 12, 10, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
    (C1, 10, 8, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
    (C1, 8, 6, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
    (C1, 6, 4, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
    (C1, 4, 2, 5, 1, 0, 0), (ACT, orc.A_LEAKYHARDTANH, 0, 0, 0, 0, 0.01),
    (FC, 80, 1, 0, 0, 0, 0), (ACT, orc.A_SOFTSIGN, 0, 0, 0, 0, 0)])
