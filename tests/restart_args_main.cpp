// Test driver of the restart argument rules (csrst_amd/host/restart.hpp):
//   restart_args_main RECORDS [ARGS...]
// parses ARGS as the driver does, applies the last record's attributes and
// prints the resulting arguments, one "name value" line each.
#include <cstdio>
#include <vector>

#include "../csrst_amd/host/restart.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    try {
        std::vector<char *> args = {argv[0]};
        for (int k = 2; k < argc; ++k) args.push_back(argv[k]);
        csrst::Commands c = csrst::Commands::parse((int)args.size(), args.data());
        csrst::restart::apply(c, csrst::ckpt::read_all(argv[1]).back());
        for (const csrst::Commands::Opt &o : c.options()) {
            std::printf("%s ", o.longName);
            switch (o.kind) {
            case 's': std::printf("%s\n", ((const std::string *)o.dst)->c_str()); break;
            case 'd': std::printf("%.17g\n", *(const double *)o.dst); break;
            case 'i': std::printf("%d\n", *(const int *)o.dst); break;
            default: std::printf("%d\n", *(const bool *)o.dst ? 1 : 0); break;
            }
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
