// accp_error.h -- error reporting shared by the library's host sources: the
// message becomes accp_last_error() of the calling thread.
#pragma once

namespace accp {
int set_error(int code, const char *msg);
}
