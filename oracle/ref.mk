# Builds the reference DIPlib library straight from its sources under
# /root/reference (read-only) into oracle/_ref/, with g++ only: no CMake, no
# generated files, no third-party libraries (pocketfft and Eigen are header-only
# and vendored in the reference tree).  Flags follow the reference's Release
# configuration (-O3 -DNDEBUG, no -march, so no FMA contraction on x86-64).
#
# Result: oracle/_ref/libdipref.so, which wraps dip::Gauss behind the C entry
# points in oracle/ref_driver.cpp.  It is test infrastructure only (golden
# fixtures and the CPU baseline); nothing in the product links it.
#
#   make -f oracle/ref.mk -j8
REF      ?= /root/reference
OUT      := oracle/_ref
CXX      ?= g++
DEFS     := -DDIP_CONFIG_DIP_IS_STATIC -DDIP_CONFIG_HAS_PRETTY_FUNCTION \
            -DDIP_CONFIG_ENABLE_STACK_TRACE -DDIP_CONFIG_ENABLE_UNICODE \
            -DDIP_VERSION_STRING='"3.6.1"' -DDIP_COPYRIGHT_YEAR='"2025"' \
            -DDIP_DEBUG_VERSION=0 -DEIGEN_MPL2_ONLY -DEIGEN_DONT_PARALLELIZE
INCS     := -I$(REF)/include -I$(REF)/dependencies/eigen3 -I$(REF)/dependencies/pocketfft
CXXFLAGS := -std=c++14 -O3 -DNDEBUG -fPIC -fopenmp -w $(DEFS) $(INCS)

SRCS := $(filter-out $(REF)/src/library/unit_tests.cpp,$(wildcard $(REF)/src/*/*.cpp))
OBJS := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(SRCS))

all: $(OUT)/libdipref.so

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libdip_static.a: $(OBJS)
	rm -f $@
	ar rcs $@ $^

$(OUT)/libdipref.so: oracle/ref_driver.cpp $(OUT)/libdip_static.a
	$(CXX) $(CXXFLAGS) -shared -o $@ oracle/ref_driver.cpp $(OUT)/libdip_static.a -fopenmp

.PHONY: all
