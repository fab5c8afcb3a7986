// SelectorFile.hpp -- reads schedule-selector weights for `grudon --selector`.
//
// The file holds the network's BF16 weights as little-endian uint16 values
// (grudon_amd/selector_bf16.bin: 5,508 of them, in the order grd_selector_load
// takes).  Decoded byte by byte, so the result does not depend on the host's
// byte order.
#pragma once

#include <cstdint>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

inline std::vector<uint16_t> read_selector_weights(const std::string &path)
{
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open selector weights " + path);
  const std::vector<char> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (in.bad()) throw std::runtime_error("cannot read selector weights " + path);
  if (bytes.size() % 2) throw std::runtime_error("selector weights " + path + ": odd byte count");
  std::vector<uint16_t> w(bytes.size() / 2);
  for (size_t i = 0; i < w.size(); i++)
    w[i] = (uint16_t)((uint16_t)(uint8_t)bytes[2 * i] | ((uint16_t)(uint8_t)bytes[2 * i + 1] << 8));
  return w;
}
