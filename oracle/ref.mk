# Builds the reference's own input parsers (plain C++17, no CUDA/HDF5) with our driver into
# oracle/_ref/ref_parse.  Reads sources where they lie under /root/reference; copies nothing.
# The reference's search path itself is CUDA + cuFFT and is not buildable here.
REF ?= /root/reference
OUT := _ref

all: $(OUT)/ref_parse

$(OUT)/ref_parse: ref_parse_driver.cpp $(REF)/src/EMReader/DataReader2.cpp
	@mkdir -p $(OUT)
	g++ -O1 -std=c++17 -I$(REF)/include/EMReader -o $@ ref_parse_driver.cpp $(REF)/src/EMReader/DataReader2.cpp

.PHONY: all
