// ReadConfig.hpp -- the reference's runtime configuration reader
// (src/ReadConfig.hpp), implemented in libvvm.so: key=value lines, '#'
// starts a comment, blanks around keys and values are dropped.
#pragma once

#include <map>
#include <string>

std::map<std::string, std::string> vvm_read_config(const std::string &filename);
