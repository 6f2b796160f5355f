// thcpp_amd.x input.json -- the reference's driver (src/main.cpp:18-77) at one
// rank on one GPU.
#include <cstdlib>
#include <iostream>
#include <stdexcept>

#include "thcpp.h"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cout << "Usage: thcpp_amd.x input.json" << std::endl;
    return EXIT_FAILURE;
  }
  try {
    return thcpp_main(argv[1]);
  } catch (const std::exception& e) {
    std::cerr << "thcpp_amd.x: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}
