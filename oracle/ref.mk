# Builds the reference CPU tick (libbiocore_gcc.a's seven sources) into oracle/_ref/.
# Test infrastructure only: the product never links or loads anything built here.
# Flags mirror the reference's own GCC recipe (CMakeLists.txt:47-84): g++ -std=c++17 -O2 -fPIC.
REF ?= /root/reference
OUT := $(dir $(lastword $(MAKEFILE_LIST)))_ref
SRCS := hh_biological receptors_biological bion_loader_host plugin_manager simulation_host consolidation dispatch_generated
OBJS := $(addprefix $(OUT)/,$(addsuffix .o,$(SRCS)))
CXXFLAGS := -std=c++17 -O2 -fPIC -w -I$(REF)/include -I$(REF)/src

all: $(OUT)/libbiocore_ref.so

$(OUT)/%.o: $(REF)/src/%.cpp
	@mkdir -p $(OUT)
	g++ $(CXXFLAGS) -c $< -o $@

$(OUT)/libbiocore_ref.so: $(OBJS)
	g++ -shared -o $@ $(OBJS) -lpthread -ldl

clean:
	rm -f $(OBJS) $(OUT)/libbiocore_ref.so
.PHONY: all clean
