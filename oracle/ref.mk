# Builds the reference's CPU library (libqubitserf, CPU backend only) from its own
# sources where they lie, into oracle/_ref/. Test infrastructure only: the result is a
# checker for the oracle restatement and the CPU baseline, never part of the product.
#
#   make -f oracle/ref.mk REF=/root/reference
#
# The reference's CMake build is not used; its CUDA and Metal translation units are not
# compiled (they need nvcc / Apple toolchains). Leon (codeaut/leon.cpp) is not on the
# enumeration path and is left out.
REF ?= /root/reference
OUT ?= oracle/_ref
CXX ?= g++
CXXFLAGS ?= -O3 -DNDEBUG -std=c++17 -fPIC -pthread

SRCS = \
  $(REF)/cpp/qsf/gf2.cpp \
  $(REF)/cpp/qsf/backend_cpu.cpp \
  $(REF)/cpp/distfind/css.cpp \
  $(REF)/cpp/distfind/stab.cpp \
  $(REF)/cpp/distfind/bz.cpp \
  $(REF)/cpp/distfind/mitm.cpp \
  $(REF)/cpp/distfind/cc.cpp \
  $(REF)/cpp/distfind/op_weight.cpp \
  $(REF)/cpp/distfind/capi.cpp \
  $(REF)/cpp/codeaut/bz.cpp \
  $(REF)/cpp/codeaut/capi.cpp

OBJS = $(patsubst $(REF)/cpp/%.cpp,$(OUT)/obj/%.o,$(SRCS))
SHIM = $(OUT)/obj/ref_shim.o

all: $(OUT)/libqubitserf_ref.so $(OUT)/cpu_enum_ref $(OUT)/libqubitserf_ref_hip.so

$(OUT)/obj/%.o: $(REF)/cpp/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -I$(REF)/cpp -I$(REF)/cpp/codeaut -c $< -o $@

# ref_shim.cpp (ours) exposes the reference's CPU enumerator for single-level timing
$(SHIM): oracle/ref_shim.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -I$(REF)/cpp -c $< -o $@

$(OUT)/libqubitserf_ref.so: $(OBJS) $(SHIM)
	$(CXX) $(CXXFLAGS) -shared $(OBJS) $(SHIM) -o $@

# The reference's own randomized-plan self-check of its CPU enumerator.
$(OUT)/cpu_enum_ref: $(REF)/cpp/distfind/tests/cpu_enum_ref.cpp $(OUT)/libqubitserf_ref.so
	$(CXX) $(CXXFLAGS) -I$(REF)/cpp $< $(OBJS) -o $@

# The reference with the gfx950 plugin in its accelerator slot: the same sources, with
# backend_cpu.cpp built -DQSF_CUDA (its cuda_backend() stub left out), the plugin
# qubitserf_amd/plugin/backend_hip.cpp, and ref_hip_slot.cpp (cuda_backend() = hip_backend()),
# linked against the engine library. -Bsymbolic keeps the reference's own distfind_* /
# codeaut_* calls inside this library, apart from the engine's same-named exports.
HIPLIB ?= qubitserf_amd/_lib
HIP_OBJS = $(filter-out $(OUT)/obj/qsf/backend_cpu.o,$(OBJS)) $(OUT)/obj/hip/backend_cpu.o \
           $(OUT)/obj/hip/backend_hip.o $(OUT)/obj/hip/ref_hip_slot.o

$(OUT)/obj/hip/backend_cpu.o: $(REF)/cpp/qsf/backend_cpu.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -DQSF_CUDA -I$(REF)/cpp -c $< -o $@

$(OUT)/obj/hip/backend_hip.o: qubitserf_amd/plugin/backend_hip.cpp include/qsf_hip.h
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -Wall -I$(REF)/cpp -Iinclude -c $< -o $@

$(OUT)/obj/hip/ref_hip_slot.o: oracle/ref_hip_slot.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -I$(REF)/cpp -c $< -o $@

$(OUT)/libqubitserf_ref_hip.so: $(HIP_OBJS) $(HIPLIB)/libqubitserf_amd.so
	$(CXX) $(CXXFLAGS) -shared -Wl,-Bsymbolic $(HIP_OBJS) -o $@ -L$(HIPLIB) -lqubitserf_amd \
	    -Wl,-rpath,'$$ORIGIN/../../qubitserf_amd/_lib'

clean:
	rm -rf $(OUT)

.PHONY: all clean
