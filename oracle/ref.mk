# Builds oracle/_ref/ref_driver: the reference's own CPU propagation/analysis code,
# compiled from its sources where they lie (read-only) plus ref_driver.cpp.
# Test infrastructure only; outputs go to oracle/_ref/ and nowhere else.
#
# Only the translation units on the propagation path are compiled (analysis/{focal,
# throughfocus,chromatic,mtf,field,tolerance}.cpp, design/{metalens,phase_profile,achromatic,
# pb_achromatic}.cpp, and the PGM writer io/image.cpp).  They need
# no Eigen, RCWA or CUDA code for the functions the driver calls; the linker's
# --gc-sections drops the unreferenced functions of those units (for example
# the library builders that would pull in the RCWA solver).  No stand-in for
# any missing header, library or generated file is used: if the reference tree
# is absent this makefile fails and the tests that need it skip.
REF     ?= /root/reference
SRC     := $(REF)/src/celeris
CXX     ?= g++
# The reference's Release flags: -O3 (CMakeLists.txt:12-14), -mavx2 -mfma (:35-47), C++20.
FLAGS   := -std=c++20 -O3 -mavx2 -mfma -ffunction-sections -fdata-sections -I$(REF)/src
OUT     := _ref
UNITS   := analysis/focal analysis/throughfocus analysis/chromatic analysis/mtf analysis/field \
           analysis/tolerance \
           design/metalens design/phase_profile design/achromatic design/pb_achromatic io/image
OBJS    := $(addprefix $(OUT)/,$(addsuffix .o,$(notdir $(UNITS))))

$(OUT)/ref_driver: $(OBJS) $(OUT)/ref_driver.o
	$(CXX) -Wl,--gc-sections -o $@ $^ -pthread

$(OUT)/ref_driver.o: ref_driver.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<

$(OUT)/focal.o: $(SRC)/analysis/focal.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/throughfocus.o: $(SRC)/analysis/throughfocus.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/chromatic.o: $(SRC)/analysis/chromatic.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/mtf.o: $(SRC)/analysis/mtf.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/field.o: $(SRC)/analysis/field.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/metalens.o: $(SRC)/design/metalens.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/phase_profile.o: $(SRC)/design/phase_profile.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/tolerance.o: $(SRC)/analysis/tolerance.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/achromatic.o: $(SRC)/design/achromatic.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/pb_achromatic.o: $(SRC)/design/pb_achromatic.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
$(OUT)/image.o: $(SRC)/io/image.cpp
	@mkdir -p $(OUT)
	$(CXX) $(FLAGS) -c -o $@ $<
