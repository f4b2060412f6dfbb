// Runs the library's field and fold routines (linac_amd/csrc/field.hpp,
// modp_gemm.hpp) on the host, for the CPU exactness tests: one command per
// input line, one result per output line.
//   F32 p cval s0..s6      fold_sub        (pinv = 1 / p in double, as the kernels)
//   K32 p cval s0..s6      fold_sub_k32
//   F64 p cval s0..s14     fold_sub64
//   U32 p h x a y          FieldU32::ff_update   (h x - a y) 2^-32 mod p
//   U64 p h x a y          FieldU64::ff_update   (h x - a y) 2^-64 mod p
//   M32 p a b / M64 p a b  mul
//   I32 p h / I64 p h      inverse
//   B32 p h / B64 p h      binary_inverse
//   E32 p h                euclid_inverse_u32
//   X61 h                  m61_inverse (p = 2^61 - 1)
//   D32 c / D64 c          balanced base-256 digits of a centred value
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "field.hpp"
#include "modp_gemm.hpp"

using namespace linac;

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string op;
        in >> op;
        if (op == "F32" || op == "K32") {
            uint64_t p, c;
            int32_t s[7];
            in >> p >> c;
            for (auto& v : s) in >> v;
            const double pd = (double)p, pinv = 1.0 / pd;
            const uint32_t r = op == "F32" ? fold_sub((uint32_t)c, s, pd, pinv) : fold_sub_k32((uint32_t)c, s, pd, pinv);
            std::cout << r << "\n";
        } else if (op == "F64") {
            uint64_t p, c;
            int32_t s[15];
            in >> p >> c;
            for (auto& v : s) in >> v;
            std::cout << fold_sub64(c, s, p, 1.0 / (double)p) << "\n";
        } else if (op == "B32" || op == "B64") {
            uint64_t p, a;
            in >> p >> a;
            if (op == "B32") std::cout << binary_inverse<uint32_t>((uint32_t)a, (uint32_t)p) << "\n";
            else std::cout << binary_inverse<uint64_t>(a, p) << "\n";
        } else if (op == "X61") {
            uint64_t a;
            in >> a;
            std::cout << m61_inverse(a) << "\n";
        } else if (op == "E32") {
            uint64_t p, a;
            in >> p >> a;
            std::cout << euclid_inverse_u32((uint32_t)a, (uint32_t)p) << "\n";
        } else if (op == "U32" || op == "M32" || op == "I32") {
            uint64_t p, a, b, c, d;
            in >> p >> a;
            const FieldU32 f = make_field_u32((uint32_t)p);
            if (op == "U32") {
                in >> b >> c >> d;
                std::cout << f.ff_update((uint32_t)a, (uint32_t)b, (uint32_t)c, (uint32_t)d) << "\n";
            } else if (op == "M32") {
                in >> b;
                std::cout << f.mul((uint32_t)a, (uint32_t)b) << "\n";
            } else {
                std::cout << f.inverse((uint32_t)a) << "\n";
            }
        } else if (op == "U64" || op == "M64" || op == "I64") {
            uint64_t p, a, b, c, d;
            in >> p >> a;
            const FieldU64 f = make_field_u64(p);
            if (op == "U64") {
                in >> b >> c >> d;
                std::cout << f.ff_update(a, b, c, d) << "\n";
            } else if (op == "M64") {
                in >> b;
                std::cout << f.mul(a, b) << "\n";
            } else {
                std::cout << f.inverse(a) << "\n";
            }
        } else if (op == "D32") {
            int64_t c;
            in >> c;
            int8_t d[4];
            balanced_digits((int32_t)c, d);
            std::cout << (int)d[0] << " " << (int)d[1] << " " << (int)d[2] << " " << (int)d[3] << "\n";
        } else if (op == "D64") {
            int64_t c;
            in >> c;
            int8_t d[8];
            balanced_digits(c, d);
            for (int i = 0; i < 8; ++i) std::cout << (int)d[i] << (i < 7 ? " " : "\n");
        } else if (!op.empty()) {
            std::cerr << "unknown command " << op << "\n";
            return 2;
        }
    }
    return 0;
}
