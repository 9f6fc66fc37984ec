// The drop-in's Editops / Opcodes algebra on the host (no GPU call): reads cases from
// stdin, one per line, and prints each result line, for tests/test_editops_algebra.py to
// compare with the reference's.  Case: op a b c src_len dest_len n (type src dest)*n
// nsub (type src dest)*nsub.  op as in oracle/ref_shim.cpp's ref_editops_algebra.
// Result: "err" or "src_len dest_len count" then the triples (editops) or 5-tuples (opcodes).
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "rapidfuzz_amd/rapidfuzz_amd.hpp"

namespace rfa = rapidfuzz_amd;

static rfa::Editops read_ops(std::istream& in, size_t src_len, size_t dest_len)
{
    size_t n = 0;
    in >> n;
    rfa::Editops e;
    for (size_t i = 0; i < n; ++i) {
        int t;
        size_t s, d;
        in >> t >> s >> d;
        e.emplace_back(static_cast<rfa::EditType>(t), s, d);
    }
    e.set_src_len(src_len);
    e.set_dest_len(dest_len);
    return e;
}

static void put(std::ostream& out, const rfa::Editops& r)
{
    out << r.get_src_len() << ' ' << r.get_dest_len() << ' ' << r.size();
    for (const rfa::EditOp& o : r) out << ' ' << static_cast<int>(o.type) << ' ' << o.src_pos << ' ' << o.dest_pos;
    out << '\n';
}

static void put(std::ostream& out, const rfa::Opcodes& r)
{
    out << r.get_src_len() << ' ' << r.get_dest_len() << ' ' << r.size();
    for (const rfa::Opcode& o : r)
        out << ' ' << static_cast<int>(o.type) << ' ' << o.src_begin << ' ' << o.src_end << ' ' << o.dest_begin << ' '
            << o.dest_end;
    out << '\n';
}

int main()
{
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) continue;
        std::istringstream in(line);
        int op, a, b, c;
        size_t sl, dl;
        in >> op >> a >> b >> c >> sl >> dl;
        rfa::Editops e = read_ops(in, sl, dl);
        rfa::Editops sub = read_ops(in, 0, 0);
        try {
            switch (op) {
            case 0: put(std::cout, e.slice(a, b, c)); break;
            case 1:
                e.remove_slice(a, b, c);
                put(std::cout, e);
                break;
            case 2: put(std::cout, e.reverse()); break;
            case 3: put(std::cout, e.inverse()); break;
            case 4: put(std::cout, e.remove_subsequence(sub)); break;
            case 5: put(std::cout, rfa::Opcodes(e).slice(a, b, c)); break;
            case 6: put(std::cout, rfa::Opcodes(e).inverse()); break;
            case 7: put(std::cout, rfa::Opcodes(e).reverse()); break;
            case 8: put(std::cout, rfa::Editops(rfa::Opcodes(e).slice(a, b, c))); break;
            default: std::cout << "err\n";
            }
        }
        catch (const std::invalid_argument&) {
            std::cout << "err\n";
        }
    }
    return 0;
}
