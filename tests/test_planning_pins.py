"""Hand-derived integer cases for every branch of the reference's planning, checked on both restatements (the
product's C++ planning through c_coil, and the oracle's C planning).

Each case follows the source step by step; the arithmetic is written beside it. Constants (PrecisionGlobalVars.h):
base increments 10 (CPU and GPU), primary measure modifiers sqrt 2 (1.414213562), secondary modifiers 1, GPU cap 80,
Legendre order cap 100; a THIN/FILAMENT thickness and a FLAT/FILAMENT length below 1e-7 R are clamped to 1e-25 R
(Coil.cxx:27-28), which makes their measures ~1e-13. Rounding is std::round (half away from zero); the total is
truncated, int(total * 2^(pf - 1)); the spill divides the total by 80 and rounds up.

Measures (CoilPairArguments.cxx:56-62): primary m0 = sqrt2 sqrt(pi (R + a/2)), m1 = sqrt2 sqrt(a), m2 = sqrt2 sqrt(b);
secondary m3 = sqrt(2 pi (R + a/2)), m4 = sqrt(a), m5 = sqrt(b). balanceIncrements (:20-46): linear =
(prod m / total)^(1/k) over the k components, layer = round(m / linear).
"""
import pytest

import oracle as O


@pytest.fixture(scope="module")
def cc():
    import c_coil
    return c_coil


def _args6(a):
    return [a.angular_block_count, a.thickness_block_count, a.length_block_count, a.angular_increment_count,
            a.thickness_increment_count, a.length_increment_count]


# (primary R, a, b), (secondary R, a, b), precision factor, GPU method, z-axis case, pure GPU, expected 12 ints
PAIR_CASES = [
    # CPU, general, RECTANGULAR-RECTANGULAR (:84-90, :125-131): components 0, 1, 3, 4, 5, total 10^5 * 2^4 = 1.6e6;
    # m = 0.970813, 0.447214, 1.188998, 0.223607, 0.547723, linear 0.0330637 -> 29.36, 13.53, 35.96, 6.76, 16.57
    # -> 29, 14, 36, 7, 17; the primary length is not a component: round(29 * 0.447214 / 0.970813) = round(13.36) = 13
    ((0.1, 0.1, 0.1), (0.2, 0.05, 0.3), 5.0, False, False, False, [1, 1, 1, 29, 14, 13, 1, 1, 1, 36, 7, 17]),
    # GPU (not pure), general, RECTANGULAR-THIN (:132-137): components 0, 1, 3, 5, total 10^4 * 16 = 160,000;
    # m = 0.831355, 0.2, 0.970813, 0.447214, linear 0.0259171 -> 32.08, 7.72, 37.46, 17.26 -> 32, 8, 37, 17; no
    # layer above 80; fill: primary length round(32 * 0.316228 / 0.831355) = round(12.17) = 12, secondary
    # thickness round(32 * 1.2e-13 / 0.83) = 0 -> 1
    ((0.1, 0.02, 0.05), (0.15, 0.0, 0.2), 5.0, True, False, False, [1, 1, 1, 32, 8, 12, 1, 1, 1, 37, 1, 17]),
    # GPU pure, general, FLAT-RECTANGULAR at PF 6: components 0, 1, 3, 4, 5, total 10^5 * 32 = 3.2e6; m = 1.482941,
    # 0.447214, 0.613996, 0.141421, 0.2, linear 0.0204758 -> 72.42, 21.84, 29.99, 6.91, 9.77 -> 72, 22, 30, 7, 10;
    # the primary length fill round(72 * 2.4e-13 / 1.48) = 0 -> 1
    ((0.3, 0.1, 0.0), (0.05, 0.02, 0.04), 6.0, True, False, True, [1, 1, 1, 72, 22, 1, 1, 1, 1, 30, 7, 10]),
    # CPU, z-axis case, FILAMENT-RECTANGULAR (:107-113): components 0 and 4 only, total 10 * 10 * 16 = 1,600;
    # m0 = 0.792665, m4 = 0.316228, linear 0.0125166 -> 63.33, 25.26 -> 63, 25; fill (:210-220): primary thickness and
    # length 0 -> 1, secondary angle round(63 * 1.253314 / 0.792665) = round(99.61) = 100, secondary length
    # round(63 * 0.223607 / 0.792665) = round(17.77) = 18; 100 increments are one block (:224-233)
    ((0.1, 0.0, 0.0), (0.2, 0.1, 0.05), 5.0, False, True, False, [1, 1, 1, 63, 1, 1, 1, 1, 1, 100, 25, 18]),
    # GPU, z-axis case, THIN-THIN (:114-117): the angle is the only component, total 10 * 16 = 160 > 80: spill
    # (:179) -> layer 0 = 80 and no component left; fill from 80: primary length round(80 * 0.632456 / 0.792665) =
    # round(63.83) = 64, secondary angle round(80 * 1.120998 / 0.792665) = round(113.14) = 113, capped (>= 80) to 80,
    # secondary length round(80 * 0.223607 / 0.792665) = round(22.57) = 23
    ((0.1, 0.0, 0.2), (0.2, 0.0, 0.05), 5.0, True, True, False, [1, 1, 1, 80, 1, 64, 1, 1, 1, 80, 1, 23]),
    # GPU (not pure), general, FILAMENT-FLAT at PF 7 (:138-143, :145-150): components 0, 3, 4, total 10^3 * 64 =
    # 64,000; m = 0.560499, 1.858965, 0.547723, linear 0.0207368 -> 27.03, 89.65, 26.41 -> 27, 90, 26; layer 3 is a
    # secondary layer, so it does not spill (first < 3 fails): the final cap makes it 80 and the total is not
    # redistributed
    ((0.05, 0.0, 0.0), (0.4, 0.3, 0.0), 7.0, True, False, False, [1, 1, 1, 27, 1, 1, 1, 1, 1, 80, 26, 1]),
    # the same pair with pure GPU: layer 3 spills (pureGPU), total ceil(64000 / 80) = 800 over components 0 and 4:
    # linear (0.560499 * 0.547723 / 800)^(1/2) = 0.0195895 -> 28.61, 27.96 -> 29, 28
    ((0.05, 0.0, 0.0), (0.4, 0.3, 0.0), 7.0, True, False, True, [1, 1, 1, 29, 1, 1, 1, 1, 1, 80, 28, 1]),
    # CPU, z-axis case, THIN-THIN at PF 8: total 10 * 128 = 1,280, the angle alone: layer 0 = 1280; fill: primary
    # length round(1280 * 0.632456 / 0.792665) = round(1021.29) = 1021, secondary angle round(1810.19) = 1810,
    # secondary length round(361.08) = 361; blocks (l - 1) / 100 + 1 = 13, 11, 19, 4 with l / blocks + 1 =
    # 99, 93, 96, 91 increments
    ((0.1, 0.0, 0.2), (0.2, 0.0, 0.05), 8.0, False, True, False, [13, 1, 11, 99, 1, 93, 19, 1, 4, 96, 1, 91]),
    # GPU (not pure), general, RECTANGULAR-FILAMENT: components 0, 1, 3, total 10^3 * 16 = 16,000; m = 1.781294,
    # 0.141421, 0.792665, linear 0.0231957 -> 76.79, 6.10, 34.17 -> 77, 6, 34 (the large primary angle stays below
    # 80); primary length round(77 * 0.141421 / 1.781294) = round(6.11) = 6
    ((0.5, 0.01, 0.01), (0.1, 0.0, 0.0), 5.0, True, False, False, [1, 1, 1, 77, 6, 6, 1, 1, 1, 34, 1, 1]),
]

# (R, a, b), precision factor, rule ("cpu", "gpu": calculatePrecisionArguments :69-185; "sec":
# getSecondaryCoilPrecisionArgumentsGPU :188-288), expected 6 ints; measures m0, m1, m2 as for the primary above
COIL_CASES = [
    # secondary GPU, RECTANGULAR (:220-225): components 0, 1, 2, total 1000 * 16 = 16,000; linear 0.0130582 ->
    # 136.41, 10.83, 10.83: the angle spills to 80, total ceil(16000 / 80) = 200 over 1 and 2: linear 0.01 ->
    # 14.14, 14.14 -> 14, 14
    ((0.5, 0.01, 0.01), 5.0, "sec", [1, 1, 1, 80, 14, 14]),
    # secondary GPU, FLAT at PF 6 (:232-237): components 0, 1, total 100 * 32 = 3,200; m = 1.188998, 0.316228,
    # linear 0.0108397 -> 109.69, 29.17: the angle spills, total ceil(3200 / 80) = 40, the thickness alone takes it;
    # the length fill round(80 * 2e-13 / 1.19) = 0 -> 1
    ((0.2, 0.05, 0.0), 6.0, "sec", [1, 1, 1, 80, 40, 1]),
    # secondary GPU, FILAMENT (:238-241): the angle alone, total 10 * 16 = 160 > 80: spill, no component left
    ((0.1, 0.0, 0.0), 5.0, "sec", [1, 1, 1, 80, 1, 1]),
    # secondary GPU, THIN at PF 9 (:226-231): components 0, 2, total 100 * 256 = 25,600; m = 0.560499, 1.414214,
    # linear 0.00556449 -> 100.73, 254.15: the angle spills (first above 80 in order), total 320; the length alone is
    # 320 > 80 and spills too, total 4 with no component left
    ((0.05, 0.0, 1.0), 9.0, "sec", [1, 1, 1, 80, 1, 80]),
    # coil CPU, RECTANGULAR at PF 12: components 0, 1, total 100 * 2048 = 204,800; m = 0.970813, 0.447214, linear
    # 0.001456 -> 666.77, 307.15 -> 667, 307; length fill round(667 * 0.447214 / 0.970813) = round(307.26) = 307;
    # blocks 7, 4, 4 with 667 / 7 + 1 = 96, 307 / 4 + 1 = 77 increments
    ((0.1, 0.1, 0.1), 12.0, "cpu", [7, 4, 4, 96, 77, 77]),
    # coil CPU, FILAMENT: the angle alone, 160 (no cap on the CPU): 2 blocks of 160 / 2 + 1 = 81
    ((0.1, 0.0, 0.0), 5.0, "cpu", [2, 1, 1, 81, 1, 1]),
    # coil GPU, FILAMENT: 160 > 80 spills to 80
    ((0.1, 0.0, 0.0), 5.0, "gpu", [1, 1, 1, 80, 1, 1]),
    # coil GPU, THIN at PF 3: the angle alone, total 10 * 4 = 40; length fill round(40 * 1.264911 / 0.354491) =
    # round(142.73) = 143, capped (> 80) to 80
    ((0.02, 0.0, 0.8), 3.0, "gpu", [1, 1, 1, 40, 1, 80]),
]


@pytest.mark.parametrize("p,s,pf,gpu,z_axis,pure,expected", PAIR_CASES)
def test_pair_planning_by_hand(cc, p, s, pf, gpu, z_axis, pure, expected):
    C = cc.coil
    method = C.GPU if gpu else C.CPU_ST
    a = C.CoilPairArguments.get_appropriate_coil_pair_arguments(C.Coil(*p, 100), C.Coil(*s, 100),
                                                               C.PrecisionFactor(pf), method, z_axis, pure)
    assert _args6(a.primary_precision) + _args6(a.secondary_precision) == expected
    assert [int(x) for x in O.pair_precision_args(O.make_coil(*p, 100), O.make_coil(*s, 100), pf, gpu, z_axis,
                                                  pure)] == expected


@pytest.mark.parametrize("g,pf,rule,expected", COIL_CASES)
def test_coil_planning_by_hand(cc, g, pf, rule, expected):
    C = cc.coil
    c, oc, P = C.Coil(*g, 100), O.make_coil(*g, 100), C.PrecisionFactor(pf)
    if rule == "sec":
        got = _args6(C.PrecisionArguments.get_secondary_coil_precision_arguments_GPU(c, P))
        ref = O.secondary_precision_args_gpu(oc, pf)
    elif rule == "gpu":
        got = _args6(C.PrecisionArguments.get_coil_precision_arguments_GPU(c, P))
        ref = O.coil_precision_args(oc, pf, True)
    else:
        got = _args6(C.PrecisionArguments.get_coil_precision_arguments_CPU(c, P))
        ref = O.coil_precision_args(oc, pf, False)
    assert got == expected
    assert [int(x) for x in ref] == expected
