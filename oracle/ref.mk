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

all: $(OUT)/libqubitserf_ref.so $(OUT)/cpu_enum_ref

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

clean:
	rm -rf $(OUT)

.PHONY: all clean
