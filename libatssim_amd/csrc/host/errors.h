// errors.h -- thread-local last-error string behind lats_last_error().
#pragma once
#include <string>

namespace lats {
void set_error(const std::string& msg);
}
