"""The rocprofv3 summariser's kernel -> operator mapping (CPU only)."""
from heterogeneous_mhd_amd.profiling import op_of, short_name

ANON = "void mhd::(anonymous namespace)::"


def test_single_box_and_ensemble_kernels_are_told_apart():
    cases = {
        ANON + "transpose_kernel(mhd::(anonymous namespace)::PairArgs, int, int)": "transpose (u and b)",
        ANON + "ens_transpose_kernel<5>(float*, float const*)": "ensemble transpose (u)",
        ANON + "ens_transpose_kernel<3>(float*, float const*)": "ensemble transpose (b)",
        ANON + "cfl_fold_kernel(int, mhd::CflOut)": "calcfl fold",
        ANON + "ens_cfl_fold_kernel(int)": "ensemble calcfl fold",
        ANON + "advect_kernel<512, 2, true, 512, true>(float*)": "advectbyzx fast",
        ANON + "ens_advect_kernel<256, 2, 64, true, true>(float*)": "ensemble advectbyzx",
        ANON + "fluidx_kernel<512, 2, false, true, 512, true>(float*)": "fluidx fast",
        ANON + "ens_fluid_kernel<256, 2, 64, false, true, true>(float*)": "ensemble fluidx",
        ANON + "advect_fluid_kernel<512, 2, true, true, 512, true, 1>(float*)":
            "advect_fluidx+calcfl fast",
        ANON + "advect_fluid_kernel<512, 2, false, true, 512, true, 2>(float*)": "advect2_fluidx fast",
        ANON + "advect_transpose_kernel<512, true, 13>(float*)": "advect_transpose13 fast",
        ANON + "transpose_advect_kernel<512, true, 12>(float*)": "transpose12_advect fast",
        ANON + "transpose_advect_kernel<512, false, 13>(float*)": "transpose13_advect",
        ANON + "fluidx_kernel<512, 2, true, true, 512, true>(float*)": "fluidx+calcfl fast",
        ANON + "ens_advect_transpose_kernel<true, 13>(mhd::(anonymous namespace)::EnsDev)":
            "ensemble advect_transpose13",
        ANON + "ens_transpose_advect_kernel<false, 12>(mhd::(anonymous namespace)::EnsDev)":
            "ensemble transpose12_advect",
    }
    for name, op in cases.items():
        assert op_of(name) == op, (name, op_of(name), op)


def test_corrector_kernels():
    assert op_of(ANON + "conv_kernel<0, 2, 1>(unsigned short const*)") == "corr_conv hidden (GELU)"
    assert op_of(ANON + "conv_kernel<1, 2, 1>(unsigned short const*)") == "corr_conv output"
    assert op_of(ANON + "conv_kernel<2, 1, 2>(unsigned short const*)") == "corr_conv input gradient"
    assert op_of(ANON + "conv_dw_kernel(unsigned short const*)") == "corr_conv_dw"
    assert op_of(ANON + "conv_first_kernel(unsigned short const*)") == "corr_conv first layer (GELU)"
    for k in ("ysum", "ymean", "apply", "check", "decide", "commit"):
        assert op_of(ANON + f"{k}_kernel(float const*)") == "corr_apply"
    assert op_of("at::native::vectorized_elementwise_kernel<4>(int)") is None


def test_short_name_keeps_template_arguments():
    assert short_name(ANON + "conv_kernel<0, 2, 1>(unsigned short const*)") == "conv_kernel<0, 2, 1>"
    assert short_name(ANON + "transpose_kernel(mhd::(anonymous namespace)::PairArgs)") == \
        "transpose_kernel"
