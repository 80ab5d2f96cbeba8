# Builds the reference CPU program from its own sources under REF_ROOT into
# oracle/_ref/.  Test infrastructure only: the outputs are used to write golden
# fixtures and as the timed CPU baseline; no product code links them.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Only the reference's C++ sources are compiled (its CMake build is not run);
# the stub scoring backend is used because the CUDA translation unit needs
# nvcc, which this image does not have.

REF_ROOT ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
CXXFLAGS ?= -O3 -std=c++17 -fopenmp -DNDEBUG

REF_SRCS := \
  $(REF_ROOT)/src/graph_io/graph_io.cpp \
  $(REF_ROOT)/src/cost_oracle.cpp \
  $(REF_ROOT)/src/quotient_graph.cpp \
  $(REF_ROOT)/src/candidate_index.cpp \
  $(REF_ROOT)/src/scoring_cpu.cpp \
  $(REF_ROOT)/src/sweg.cpp \
  $(REF_ROOT)/src/cuda_scoring_stub.cpp

REF_OBJS := $(patsubst $(REF_ROOT)/src/%.cpp,$(OUT)/obj/%.o,$(REF_SRCS))

all: $(OUT)/beam $(OUT)/gain_harness

$(OUT)/obj/%.o: $(REF_ROOT)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -I$(REF_ROOT)/include -c $< -o $@

$(OUT)/beam: $(REF_OBJS) $(REF_ROOT)/src/main.cpp
	$(CXX) $(CXXFLAGS) -I$(REF_ROOT)/include $(REF_ROOT)/src/main.cpp $(REF_OBJS) -o $@

# Per-pair gain harness: this repository's driver (oracle/ref_gain_harness.cpp)
# linked against the reference QuotientGraph and CostOracle.
$(OUT)/gain_harness: oracle/ref_gain_harness.cpp $(REF_OBJS)
	$(CXX) $(CXXFLAGS) -I$(REF_ROOT)/include $< $(REF_OBJS) -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
