# CPU-only AddressSanitizer + UndefinedBehaviorSanitizer build of the host
# code that runs without a GPU: the library's text-format reader and writers
# (heterogeneous_mhd_amd/csrc/mhd_io.hip), its seeded host IC generator
# (mhd_ic.hip) and the oracle restatement (oracle/mhd_ref.c, FP32 and FP64),
# driven by tests/c/sanitize_harness.cpp.  Used by tests/test_sanitize.py.
#
#   make -f tests/c/sanitize.mk [OUT=dir]
#
# The HIP sources are compiled by hipcc with the sanitizers on the host pass
# only (-Xarch_host before each flag; device code is built as usual), and
# linked with ROCm's clang++, whose sanitizer runtimes they need.

ROOT   := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/../..)
OUT    ?= $(ROOT)/tests/c/_san
ROCM   ?= /opt/rocm
HIPCC  ?= $(ROCM)/bin/hipcc
CXX    := $(ROCM)/llvm/bin/clang++
CC     := $(ROCM)/llvm/bin/clang
CSRC   := $(ROOT)/heterogeneous_mhd_amd/csrc
SAN    := -fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer
HIPSAN := -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
          -Xarch_host -fno-sanitize-recover=all -Xarch_host -fno-omit-frame-pointer
HIPFLAGS := --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -ffp-contract=off \
            -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-command-line-argument \
            -I$(ROOT)/include $(HIPSAN)

.PHONY: all
all: $(OUT)/sanitize_f32 $(OUT)/sanitize_f64

$(OUT)/mhd_io.o: $(CSRC)/mhd_io.hip $(CSRC)/mhd_internal.h $(ROOT)/include/mhd_hip.h
	mkdir -p $(OUT)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(OUT)/mhd_ic.o: $(CSRC)/mhd_ic.hip $(CSRC)/mhd_internal.h
	mkdir -p $(OUT)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(OUT)/ref_%.o: $(ROOT)/oracle/mhd_ref.c
	mkdir -p $(OUT)
	$(CC) -O1 -g -ffp-contract=off $(SAN) -DREAL=$(if $(filter f32,$*),float,double) -c $< -o $@

$(OUT)/sanitize_%: $(ROOT)/tests/c/sanitize_harness.cpp $(OUT)/ref_%.o $(OUT)/mhd_io.o $(OUT)/mhd_ic.o
	$(CXX) -O1 -g -std=c++17 $(SAN) -DREAL=$(if $(filter f32,$*),float,double) -I$(ROOT)/include $^ \
	    -o $@ -L$(ROCM)/lib -lamdhip64 -Wl,-rpath,$(ROCM)/lib -lm
